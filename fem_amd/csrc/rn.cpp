// Reflective Newton (LinA/reflective_newton.c) on the device.
//
// reflective_newton (:660-711) minimises q(x) = x.H.x/2 + x.b over a box by
// trust-region steps in the scaled variables of Coleman-Li: D_i =
// sqrt|x_i - bound_i| on bounded entries, M = D H D + diag(|g| J).  Each step
// (RN_iterate, :536-627) solves M v = -D g with an inner Krylov method, takes
// the Newton step if it is inside the trust region and otherwise the minimum
// of q on the circle of radius Delta in span{D g, v} (subspace_minimum,
// :394-450, roots of dq/dtheta by GMRES_Newton.c:353-439), reflects the
// iterate into the box, bisects until the sigma_l/sigma_u tests hold and
// updates the trust radius from the achieved/predicted reduction.
//
// Device: every length-n vector (gradient, scalings, subspace basis, steps,
// iterates), the SpMVs with H and M, dots and the inner solvers (ILU(0)-PCG,
// BiCGStab + AMG, AMG_solve).  Host: the scalars that steer the method and
// the assembly of M with the reference's zero deletion, because the ILU(0)
// pattern and the Galerkin products depend on it.

#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <vector>

#include "../../include/lina_hip.h"
#include "ilu.hpp"
#include "kernels.hpp"
#include "lina_internal.hpp"
#include "solver.hpp"

namespace lina {

// power_method (linear_algebra.c:4802-4822) with A's device SpMV: v (device,
// length n) holds the start vector and ends as the last normalised iterate;
// *eig enters as the estimate the first convergence test compares with (the
// reference reads its caller's variable there) and leaves as the last
// estimate.  q = matrix_norm(A) (linear_algebra.c:3441-3450).  Each
// iteration: Prev = v, v = A Prev, eig = v.Prev, v /= |v|, stop when
// |eig - prev| / q <= eps.  tmp: n doubles, part: block partials of the
// dots, scal: 1 device double, host: 1 pinned double.  Returns the
// iteration count, 0 when max_iter is reached (the reference's return).
int power_iteration(hipStream_t st, DevMatrix& A, double q, double* v, double* tmp, double* part, double* scal,
                    double* host, int max_iter, double eps, double* eig) {
    const int n = A.n();
    auto dot = [&](const double* a, const double* b) {
        int np = dev::dot_part<double>(st, n, a, b, part);
        dev::finalize(st, np, part, scal, 0, 0);
        hip_check(hipMemcpyAsync(host, scal, sizeof(double), hipMemcpyDeviceToHost, st), "d2h");
        hip_check(hipStreamSynchronize(st), "sync");
        return host[0];
    };
    int i = 0;
    double r;
    do {
        const double prev = *eig;
        dev::copy(st, n, (const double*)v, tmp);
        if (A.stream() != st) hip_check(hipStreamSynchronize(st), "sync");
        A.launch(tmp, nullptr, v, dev::MODE_SET, nullptr, nullptr);
        if (A.stream() != st) hip_check(hipStreamSynchronize(A.stream()), "sync");
        *eig = dot(v, tmp);
        const double s = 1. / std::sqrt(dot(v, v));
        dev::scale(st, n, v, s);
        r = std::fabs(*eig - prev) / q;
        i++;
    } while (r > eps && i < max_iter);
    return i == max_iter ? 0 : i;
}

namespace {

// ------------------------------------------------------------ host math

// 1-D root search of subspace_minimum: the functions of GMRES_Newton.c with
// the reference's constants.  `rand()` in the secant fallback is replaced by
// a private linear congruential sequence (the reference's depends on the
// process's earlier rand() calls).
struct RootSearch {
    double ax, ay, axy, bx, by, delta;
    uint64_t lcg = 0x2545F4914F6CDD1DULL;

    double uniform01() {
        lcg = lcg * 6364136223846793005ULL + 1442695040888963407ULL;
        return (double)(lcg >> 11) * (1.0 / 9007199254740992.0);
    }
    // RN_derivative (reflective_newton.c:372-374)
    double F(double x) const {
        return ((std::sin(2. * x) * (-ax + ay) + std::cos(2. * x) * axy) * delta - std::sin(x) * bx +
                std::cos(x) * by) *
               delta;
    }
    // quadratic_on_circle (:368-371)
    double q(double x) const {
        double c = std::cos(x) * delta;
        double s = std::sin(x) * delta;
        return c * (c * ax + s * axy + bx) + s * (s * ay + by);
    }
    // der_F / der_poly (GMRES_Newton.c:320-334): central differences, dx = 1e-6
    double der_F(double x, int depth) const {
        const double dx = 1e-6;
        if (depth == 0) return F(x);
        return (der_F(x + dx, depth - 1) - der_F(x - dx, depth - 1)) / (2. * dx);
    }
    static double root_poly(const std::vector<double>& r, double x) {
        double res = 1.;
        for (double v : r) res *= (x - v);
        return res;
    }
    static double der_poly(const std::vector<double>& r, double x, int depth) {
        const double dx = 1e-6;
        if (depth == 0) return root_poly(r, x);
        return (der_poly(r, x + dx, depth - 1) - der_poly(r, x - dx, depth - 1)) / (2. * dx);
    }
    // F_hospital (:336-351): F deflated by the roots found so far, by
    // l'Hospital's rule where both vanish.  The depth is capped at 12 (the
    // reference loops without bound on an identically vanishing F).
    double hospital(const std::vector<double>& roots, double x) const {
        const double eps = 1e-5;
        int depth = 0;
        double dp, df;
        do {
            dp = der_poly(roots, x, depth);
            df = der_F(x, depth);
            depth++;
        } while (std::fabs(dp) < eps && std::fabs(df) < eps && depth < 12);
        return df / dp;
    }
    // Newton_1D (:353-400): secant iterations with deflation until one fails
    std::vector<double> newton(double mn, double mx, int max_iter, double eps) {
        std::vector<double> roots;
        int i;
        do {
            double local_eps = eps;
            double x0 = mn, x1 = mx;
            double f0 = hospital(roots, x0);
            double f1 = hospital(roots, x1);
            double F0 = f0 != 0 ? f0 : 1.;
            double x = 0, f = 0;
            i = 0;
            do {
                x = x1 - f1 * (x1 - x0) / (f1 - f0);
                if (x < mn || x > mx) x = mn + (mx - mn) * uniform01();
                f = hospital(roots, x);
                x0 = x1;
                x1 = x;
                f0 = f1;
                f1 = f;
                i++;
                if (i == max_iter / 2) local_eps *= 1e2;
            } while (std::fabs(f / F0) > local_eps && i < max_iter);
            if (i < max_iter) roots.push_back(x);
        } while (i < max_iter);
        return roots;
    }
    // Interval_Separation_Newton_1D (:402-439) on [0, 2 pi]
    std::vector<double> roots() {
        const int int_num = 1000, max_iter = 500;
        const double eps = 1e-6, mn = 0, mx = 2. * M_PI;
        const double d = (double)(mx - mn) / int_num;
        std::vector<double> out;
        for (int i = 0; i < int_num; i++) {
            double l = i * d;
            double r = l + d;
            double fl = F(l), fr = F(r);
            if (fl * fr < 0) {
                std::vector<double> loc = newton(l, r, max_iter, eps);
                // appended in reverse order of discovery, as the reference does
                size_t base = out.size();
                out.resize(base + loc.size());
                for (size_t j = 0; j < loc.size(); j++) out[out.size() - j - 1] = loc[j];
            } else {
                if (fl == 0) out.push_back(fl);
                if (fr == 0) out.push_back(fr);
            }
        }
        return out;
    }
};

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// LINA_RN_TIMING set: one line per Newton step with the time of its phases
bool rn_timing() {
    const char* e = getenv("LINA_RN_TIMING");  // read per call: a process may switch it
    return e && atoi(e) != 0;
}

uint64_t splitmix(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

}  // namespace

// RN_get_M_D's product D H D (linear_algebra.c:186-222 twice, then
// delete_zeros) plus insert_sparse(M, |g_i| J_i, i, i) (reflective_newton.c:
// 250-262), with the reference's rounding: M_ij = D_i (H_ij D_j), the
// diagonal accumulated onto sparse_zero's 0.0 seed.
Csr scaled_hessian(const Csr& H, const std::vector<double>& D, const std::vector<double>& G,
                   const std::vector<double>& J) {
    const int n = H.n;
    Csr M;
    M.n = n;
    M.ncols = n;
    M.rp.assign(n + 1, 0);
    M.ci.reserve(H.ci.size());
    M.v.reserve(H.v.size());
    std::vector<int> hc;
    std::vector<double> hv;
    for (int i = 0; i < n; i++) {
        // HD row i: seed (i, 0.0); a = H_ik != 0, b = D_k != 0 -> a*b
        hc.clear();
        hv.clear();
        bool seeded = false;
        for (int p = H.rp[i]; p < H.rp[i + 1]; p++) {
            const int k = H.ci[p];
            if (!seeded && k >= i) {
                if (k > i) {
                    hc.push_back(i);
                    hv.push_back(0.0);
                }
                seeded = true;
            }
            const double a = H.v[p];
            const double b = D[k];
            double val = (a != 0 && b != 0) ? a * b : 0.0;
            if (k == i) {
                hc.push_back(i);
                hv.push_back(0.0 + val);
            } else if (a != 0 && b != 0) {
                hc.push_back(k);
                hv.push_back(val);
            }
        }
        if (!seeded) {
            hc.push_back(i);
            hv.push_back(0.0);
        }
        // delete_zeros on HD's row (a row of zeros keeps (i, 0.0))
        {
            size_t w = 0;
            if (hc.size() > 1) {
                for (size_t t = 0; t < hc.size(); t++)
                    if (hv[t] != 0) {
                        hc[w] = hc[t];
                        hv[w] = hv[t];
                        w++;
                    }
                if (w == 0) {
                    hc[0] = i;
                    hv[0] = 0.0;
                    w = 1;
                }
                hc.resize(w);
                hv.resize(w);
            }
        }
        // M row i = D_i * HD row i on the seed (i, 0.0), then delete_zeros
        const int start = (int)M.ci.size();
        const double di = D[i];
        bool mseeded = false;
        for (size_t t = 0; t < hc.size(); t++) {
            const int j = hc[t];
            if (!mseeded && j >= i) {
                if (j > i) {
                    M.ci.push_back(i);
                    M.v.push_back(0.0);
                }
                mseeded = true;
            }
            const double b = hv[t];
            if (j == i) {
                M.ci.push_back(i);
                M.v.push_back(di != 0 && b != 0 ? 0.0 + di * b : 0.0);
            } else if (di != 0 && b != 0) {
                M.ci.push_back(j);
                M.v.push_back(di * b);
            }
        }
        if (!mseeded) {
            M.ci.push_back(i);
            M.v.push_back(0.0);
        }
        int len = (int)M.ci.size() - start;
        if (len > 1) {
            int w = start;
            for (int t = start; t < start + len; t++)
                if (M.v[t] != 0) {
                    M.ci[w] = M.ci[t];
                    M.v[w] = M.v[t];
                    w++;
                }
            if (w == start) {
                M.ci[w] = i;
                M.v[w] = 0.0;
                w++;
            }
            M.ci.resize(w);
            M.v.resize(w);
        }
        // insert_sparse(M, |g_i| J_i, i, i)
        if (J[i] != 0) {
            const double add = std::fabs(G[i]) * J[i];
            int pos = start, end = (int)M.ci.size();
            while (pos < end && M.ci[pos] < i) pos++;
            if (pos < end && M.ci[pos] == i) {
                M.v[pos] += add;
            } else {
                M.ci.insert(M.ci.begin() + pos, i);
                M.v.insert(M.v.begin() + pos, add);
            }
        }
        M.rp[i + 1] = (int)M.ci.size();
    }
    return M;
}

class ReflectiveNewton {
  public:
    ReflectiveNewton(const Csr& H, const double* b, int nb, const int* idx, const double* lo, const double* up,
                     const lina_rn_params& p)
        : H_(H), b_(b, b + H.n), p_(p), delta_(p.delta), seed_(p.seed) {
        n_ = H.n;
        idx_.assign(idx, idx + nb);
        lo_.assign(lo, lo + nb);
        up_.assign(up, up + nb);
        for (int k : idx_)
            if (k < 0 || k >= n_) throw std::invalid_argument("bound index out of range");
        ilu_.set_ordering(p.ilu_ordering);
        dH_.reset(new DevMatrix(H_));
        st_ = dH_->stream();
        B_.reset(new DevBuffers());
        auto alloc = [&](size_t m) {
            void* q = nullptr;
            hip_check(hipMalloc(&q, std::max<size_t>(m, 1) * sizeof(double)), "hipMalloc");
            B_->ptrs.push_back(q);
            return (double*)q;
        };
        for (double** v : {&X_, &Xo_, &G_, &D_, &J_, &V1_, &V2_, &bs_, &Bv_, &S_, &T_, &T2_}) *v = alloc(n_);
        // block partials: a row block holds >= 1 row, so n + 8192 covers H, M and the dots
        part_ = alloc((size_t)n_ + 8192);
        scal_ = alloc(16);
        lod_ = alloc(nb);
        upd_ = alloc(nb);
        void* q = nullptr;
        hip_check(hipMalloc(&q, std::max(nb, 1) * sizeof(int)), "hipMalloc");
        B_->ptrs.push_back(q);
        idxd_ = (int*)q;
        hip_check(hipMemcpyAsync(idxd_, idx_.data(), nb * sizeof(int), hipMemcpyHostToDevice, st_), "h2d");
        hip_check(hipMemcpyAsync(lod_, lo_.data(), nb * sizeof(double), hipMemcpyHostToDevice, st_), "h2d");
        hip_check(hipMemcpyAsync(upd_, up_.data(), nb * sizeof(double), hipMemcpyHostToDevice, st_), "h2d");
        hip_check(hipMemcpyAsync(bs_, b_.data(), n_ * sizeof(double), hipMemcpyHostToDevice, st_), "h2d");
        hip_check(hipHostMalloc((void**)&host_, 16 * sizeof(double)), "hipHostMalloc");
        bdev_ = alloc(n_);
        hip_check(hipMemcpyAsync(bdev_, b_.data(), n_ * sizeof(double), hipMemcpyHostToDevice, st_), "h2d");
        if (p_.solver == LINA_RN_BICGSTAB || p_.solver == LINA_RN_AMG) {
            amg_.reset(new Solver());
            amg_->setup_host(H_, p_.amg);
            amg_->upload();
        }
        if (p_.solver == LINA_RN_PCG || p_.solver == LINA_RN_GMRES) {
            // F = H's pattern plus the diagonal, for the device assembly of M
            F_.n = F_.ncols = n_;
            F_.rp.assign(n_ + 1, 0);
            std::vector<double> fhv;
            std::vector<uint8_t> fh;
            for (int i = 0; i < n_; i++) {
                bool diag = false;
                for (int q = H_.rp[i]; q < H_.rp[i + 1]; q++) {
                    const int k = H_.ci[q];
                    if (!diag && k >= i) {
                        if (k > i) {
                            F_.ci.push_back(i);
                            fhv.push_back(0.0);
                            fh.push_back(0);
                        }
                        diag = true;
                    }
                    F_.ci.push_back(k);
                    fhv.push_back(H_.v[q]);
                    fh.push_back(1);
                }
                if (!diag) {
                    F_.ci.push_back(i);
                    fhv.push_back(0.0);
                    fh.push_back(0);
                }
                F_.rp[i + 1] = (int)F_.ci.size();
            }
            frp_ = dupload(F_.rp, *B_, st_);
            fci_ = dupload(F_.ci, *B_, st_);
            fhv_ = dupload(fhv, *B_, st_);
            fh_ = dupload(fh, *B_, st_);
            mflag_ = dalloc<int>(1, *B_);
            // M's products on F's pattern (zeros where M has no entry): one
            // device matrix for every Newton step, patterns changing or not
            mvF_ = dalloc<double>(F_.ci.size(), *B_);
            F_.v.assign(F_.ci.size(), 0.0);
            dMF_.reset(new DevMatrix(F_));
            F_.v.clear();
            f_ilu_ = p_.ilu_ordering == LINA_ILU_MULTICOLOUR;
        }
        hip_check(hipStreamSynchronize(st_), "rn upload");
    }

    ~ReflectiveNewton() {
        if (st_) (void)hipStreamSynchronize(st_);
        if (host_) (void)hipHostFree(host_);
    }

    void set_delta(double d) { delta_ = d; }

    int solve(double* x, lina_rn_stats* stats) {
        const int n = n_;
        stats_ = lina_rn_stats{};
        hip_check(hipMemcpyAsync(X_, x, n * sizeof(double), hipMemcpyHostToDevice, st_), "h2d");
        double q0 = functional(X_);
        int iter = 0;
        double q = q0, dq = 0;
        std::vector<double> hD(n), hG(n), hJ(n);
        do {
            // G = b + H X (linear_map(G, 1, H, X) on a copy of b)
            dev::copy(st_, n, (const double*)bdev_, G_);
            dH_->launch(X_, nullptr, G_, dev::MODE_ADD, nullptr, nullptr);
            dev::fill<double>(st_, n, D_, 1.0);
            dev::fill<double>(st_, n, J_, 0.0);
            dev::rn_dj(st_, (int)idx_.size(), idxd_, lod_, upd_, G_, X_, D_, J_);
            const double t0 = now_s();
            // M on the device while its pattern is the cached one (the same
            // values as the host assembly), else assembled on the host
            bool on_dev = false;
            mc_current_ = false;
            if (frp_) {
                // the multicolour ILU factors M on F's pattern: no cache needed
                const bool cached = !f_ilu_ && mc_ready_ && !host_m_forced();
                dev::fill<int>(st_, 1, mflag_, 0);
                dev::rn_scaled_hessian(st_, n, frp_, fci_, fhv_, fh_, D_, G_, J_, cached ? mpos_ : nullptr,
                                       cached ? mv_ : nullptr, mvF_, mflag_);
                hip_check(hipMemcpyAsync(host_ + 8, mflag_, sizeof(int), hipMemcpyDeviceToHost, st_), "d2h");
                hip_check(hipStreamSynchronize(st_), "sync");
                int flag = 1;
                std::memcpy(&flag, host_ + 8, sizeof(int));
                on_dev = f_ilu_ || (cached && flag == 0);
                if (on_dev) mc_stale_ = true;
                dMF_->update_values_device(mvF_, (long)F_.ci.size());
            }
            if (!on_dev) {
                download(D_, hD);
                download(G_, hG);
                download(J_, hJ);
                Mc_ = scaled_hessian(H_, hD, hG, hJ);
                mc_stale_ = false;
                mc_current_ = true;
                if (frp_) cache_pattern();
            }
            t_m_ = now_s() - t0;
            iterate(on_dev);
            if (rn_timing())
                fprintf(stderr,
                        "[rn] step %d: M %.3fs (nnz %d, %s), inner %.3fs (build %.3fs, %d its), rest %.3fs\n", iter,
                        t_m_, f_ilu_ ? (int)F_.ci.size() : Mc_.nnz(),
                        f_ilu_ ? "device, F's pattern" : (on_dev ? "device" : "host"), t_inner_, t_build_,
                        last_inner_, now_s() - t0 - t_m_ - t_inner_);
            q = functional(X_);
            dq = std::fabs((q - q0) / q0);
            iter++;
            q0 = q;
        } while (iter < p_.max_iter && dq > p_.eps);
        dev::rn_snap(st_, (int)idx_.size(), idxd_, lod_, upd_, X_, 1e-3 * p_.eps);
        hip_check(hipMemcpyAsync(x, X_, n * sizeof(double), hipMemcpyDeviceToHost, st_), "d2h");
        hip_check(hipStreamSynchronize(st_), "sync");
        stats_.iterations = iter;
        stats_.dq = dq;
        stats_.q = q;
        stats_.delta = delta_;
        if (stats) *stats = stats_;
        return iter;
    }

  private:
    // ------------------------------------------------------- device helpers
    void download(const double* d, std::vector<double>& h) {
        hip_check(hipMemcpyAsync(h.data(), d, n_ * sizeof(double), hipMemcpyDeviceToHost, st_), "d2h");
        hip_check(hipStreamSynchronize(st_), "sync");
    }
    double read(int slot) {
        hip_check(hipMemcpyAsync(host_, scal_ + slot, sizeof(double), hipMemcpyDeviceToHost, st_), "d2h");
        hip_check(hipStreamSynchronize(st_), "sync");
        return host_[0];
    }
    double dot(const double* a, const double* b) {
        int np = dev::dot_part<double>(st_, n_, a, b, part_);
        dev::finalize(st_, np, part_, scal_, 0, 0);
        return read(0);
    }
    double norm(const double* a) { return std::sqrt(dot(a, a)); }
    void axpy(double* x, double f, const double* y) { dev::axpby(st_, n_, 1.0, x, f, y); }  // vector_add
    // sparse_bilinear(y, A, x) = y . (A x), A given as a DevMatrix
    double bilinear(DevMatrix& A, const double* y, const double* x) {
        sync_with(A);
        int np = A.launch(x, nullptr, T2_, dev::MODE_SET, part_, y);
        dev::finalize(A.stream(), np, part_, scal_, 1, 0);
        hip_check(hipMemcpyAsync(host_, scal_ + 1, sizeof(double), hipMemcpyDeviceToHost, A.stream()), "d2h");
        hip_check(hipStreamSynchronize(A.stream()), "sync");
        return host_[0];
    }
    void sync_with(DevMatrix& A) {
        if (A.stream() != st_) hip_check(hipStreamSynchronize(st_), "sync");
    }
    // RN_default_func: x.H.x/2 + b.x
    double functional(const double* x) {
        int np = dH_->launch(x, nullptr, T_, dev::MODE_SET, part_, x);
        dev::finalize(st_, np, part_, scal_, 2, 0);
        np = dev::dot_part<double>(st_, n_, bdev_, x, part_ + 4096);
        dev::finalize(st_, np, part_ + 4096, scal_, 3, 0);
        hip_check(hipMemcpyAsync(host_, scal_ + 2, 2 * sizeof(double), hipMemcpyDeviceToHost, st_), "d2h");
        hip_check(hipStreamSynchronize(st_), "sync");
        return host_[0] / 2. + host_[1] + p_.q_offset;
    }
    void normalize(double* v) {
        double a = norm(v);
        dev::scale(st_, n_, v, (double)1 / a);
    }

    // ------------------------------------------------------ inner solvers
    // RN_get_subspace (reflective_newton.c:294-366): V1 = D g, V2 solves
    // M V2 = -V1 from the start V2 = -V1; returns POSITIVE_DEFINITE (1) or
    // INDEFINITE (0)
    int subspace(bool on_dev, DevMatrix& dM) {
        const Csr& M = Mc_;  // values current unless on_dev (host_M() refreshes them)
        const int n = n_;
        dev::copy(st_, n, (const double*)G_, V1_);
        dev::mul(st_, n, V1_, D_);
        dev::copy(st_, n, (const double*)V1_, V2_);
        dev::scale(st_, n, V2_, -1.);
        dev::copy(st_, n, (const double*)V2_, bs_);
        hip_check(hipStreamSynchronize(st_), "sync");
        const double t_in = now_s();
        int status = 1, iter = 0;
        switch (p_.solver) {
            case LINA_RN_PCG: {
                // one ILU system for every Newton step: while M's pattern
                // stays the same only the numeric factorisation runs (device)
                const double tb = now_s();
                refactor_ilu(on_dev);
                t_build_ = now_s() - tb;
                if (!ilu_.ilu_ok()) {
                    status = 0;
                    break;
                }
                double res = 0;
                iter = ilu_.pcg_dev(V2_, bs_, p_.eps, p_.pcg_iter, LINA_STOP_REFERENCE, &res);
                status = iter < p_.pcg_iter ? 1 : 0;
                break;
            }
            case LINA_RN_GMRES: {
                // RN_UMF's slot (reflective_newton.c:312-315): the local solver
                // decides the status; FEM.c's convention y = V2, V2 = M^-1 y
                const double tb = now_s();
                refactor_ilu(on_dev);
                t_build_ = now_s() - tb;
                if (!ilu_.ilu_ok()) {
                    status = 0;
                    break;
                }
                double eta = 0;
                iter = ilu_.gmres_dev(V2_, bs_, p_.gmres_m, p_.gmres_tol, p_.gmres_max_iter, &eta);
                ilu_.apply(V2_, T_);
                hip_check(hipStreamSynchronize(ilu_.stream()), "sync");
                dev::copy(st_, n, (const double*)T_, V2_);
                status = eta <= p_.gmres_tol ? 1 : 0;
                break;
            }
            case LINA_RN_BICGSTAB: {
                const double tb = now_s();
                amg_->recompute_matrix(M);
                t_build_ = now_s() - tb;
                double res = 0;
                iter = amg_->bicgstab_dev(V2_, bs_, p_.bi_eps, p_.bi_max_iter, &res);
                hip_check(hipStreamSynchronize(amg_->stream()), "sync");
                status = iter < p_.bi_max_iter ? 1 : 0;
                break;
            }
            case LINA_RN_AMG: {
                const double tb = now_s();
                amg_->recompute_matrix(M);
                t_build_ = now_s() - tb;
                iter = amg_->amg_solve_dev(bs_, V2_);
                hip_check(hipStreamSynchronize(amg_->stream()), "sync");
                status = iter < p_.amg.amg_max_iter ? 1 : 0;
                break;
            }
            case LINA_RN_HOST: {
                // RN_UMF (reflective_newton.c:307-310): the caller's solver on
                // the host copy of M, its return value the status
                std::vector<double> hx(n), hb(n);
                download(V2_, hx);
                download(bs_, hb);
                status = p_.host_solver(p_.host_ctx, n, M.rp.data(), M.ci.data(), M.v.data(), hx.data(), hb.data()) > 0
                             ? 1
                             : 0;
                hip_check(hipMemcpyAsync(V2_, hx.data(), n * sizeof(double), hipMemcpyHostToDevice, st_), "h2d");
                hip_check(hipStreamSynchronize(st_), "sync");
                iter = 1;
                break;
            }
            default:
                throw std::invalid_argument("unknown inner solver");
        }
        stats_.inner_iterations += iter;
        t_inner_ = now_s() - t_in;
        last_inner_ = iter;
        if (status == 0) {
            stats_.neg_curvature++;
            neg_curvature(host_M(), dM);
        }
        return status;
    }

    // the ILU(0) of the step's M: the multicolour order factors M on F's
    // fixed pattern (zeros where M has no entry) from the device values,
    // built once; the natural order factors M's own pattern, from the device
    // values while it is the cached one
    void refactor_ilu(bool on_dev) {
        if (f_ilu_) {
            if (ilu_f_built_ && ilu_.refactor_device(mvF_)) return;
            Csr Fv = F_;
            Fv.v.resize(F_.ci.size());
            hip_check(hipMemcpyAsync(Fv.v.data(), mvF_, Fv.v.size() * sizeof(double), hipMemcpyDeviceToHost, st_),
                      "d2h");
            hip_check(hipStreamSynchronize(st_), "sync");
            ilu_.refactor(Fv);
            ilu_f_built_ = true;
            return;
        }
        if (!on_dev || !ilu_.refactor_device(mv_)) ilu_.refactor(host_M());
    }

    // the scaled Hessian's host copy, its values fetched after a device step
    // (assembled from D, G, J when the step's M lives on F's pattern only)
    const Csr& host_M() {
        if (f_ilu_ && !mc_current_) {
            std::vector<double> hD(n_), hG(n_), hJ(n_);
            download(D_, hD);
            download(G_, hG);
            download(J_, hJ);
            Mc_ = scaled_hessian(H_, hD, hG, hJ);
            mc_current_ = true;
            mc_stale_ = false;
            return Mc_;
        }
        if (mc_stale_) {
            hip_check(hipMemcpyAsync(Mc_.v.data(), mv_, Mc_.v.size() * sizeof(double), hipMemcpyDeviceToHost, st_),
                      "d2h");
            hip_check(hipStreamSynchronize(st_), "sync");
            mc_stale_ = false;
        }
        return Mc_;
    }

    // after a host assembly: each F position's entry in M (M's pattern lies
    // in F), for the device assembly of the next steps
    void cache_pattern() {
        mc_ready_ = false;
        std::vector<int> mpos(F_.ci.size(), -1);
        for (int i = 0; i < n_; i++) {
            int q = Mc_.rp[i];
            for (int p = F_.rp[i]; p < F_.rp[i + 1]; p++)
                if (q < Mc_.rp[i + 1] && Mc_.ci[q] == F_.ci[p]) mpos[p] = q++;
            if (q != Mc_.rp[i + 1]) return;  // an entry outside F: the host path stays
        }
        hip_check(hipStreamSynchronize(st_), "sync");
        mbufs_.reset(new DevBuffers());
        mpos_ = dupload(mpos, *mbufs_, st_);
        mv_ = dalloc<double>(std::max<size_t>(Mc_.v.size(), 1), *mbufs_);
        hip_check(hipMemcpyAsync(mv_, Mc_.v.data(), Mc_.v.size() * sizeof(double), hipMemcpyHostToDevice, st_),
                  "h2d");
        hip_check(hipStreamSynchronize(st_), "sync");
        mc_ready_ = true;
    }
    static bool host_m_forced() {
        const char* e = getenv("LINA_RN_HOST_M");
        return e && atoi(e) != 0;
    }

    // power_method on A; the first eigenvalue estimate starts at 0 (the
    // reference reads an uninitialised variable there)
    double power(DevMatrix& A, const Csr& Ah, double* v, int max_iter, double eps) {
        double q = 0;
        for (double a : Ah.v) q += std::fabs(a);
        q = q / Ah.n;  // matrix_norm (:3441-3450)
        double eig = 0;
        power_iteration(st_, A, q, v, T_, part_, scal_, host_, max_iter, eps, &eig);
        return eig;
    }
    void random_unit(double* v) {
        std::vector<double> h(n_);
        for (int i = 0; i < n_; i++) h[i] = (double)(splitmix(seed_) >> 11) * (1.0 / 9007199254740992.0);
        hip_check(hipMemcpyAsync(v, h.data(), n_ * sizeof(double), hipMemcpyHostToDevice, st_), "h2d");
        hip_check(hipStreamSynchronize(st_), "sync");
        normalize(v);
    }
    // get_neg_curvature_direction (linear_algebra.c:4849-4865): the largest
    // eigenvalue s of M, then the dominant eigenvector of M - s I
    void neg_curvature(const Csr& M, DevMatrix& dM) {
        const int eig_iter = 10000;
        const double eig_eps = 1e-6;
        random_unit(V2_);
        double shift = power(dM, M, V2_, 1000, eig_eps);
        Csr Bm = M;
        for (int i = 0; i < Bm.n; i++) {
            int pos = Bm.rp[i];
            while (pos < Bm.rp[i + 1] && Bm.ci[pos] < i) pos++;
            if (pos < Bm.rp[i + 1] && Bm.ci[pos] == i) {
                Bm.v[pos] += -shift;
            } else {
                Bm.ci.insert(Bm.ci.begin() + pos, i);
                Bm.v.insert(Bm.v.begin() + pos, -shift);
                for (int k = i + 1; k <= Bm.n; k++) Bm.rp[k]++;
            }
        }
        DevMatrix dB(Bm);
        random_unit(V2_);
        double curv = power(dB, Bm, V2_, eig_iter, eig_eps) + shift;
        last_curvature_ = curv;
    }

    // ---------------------------------------------------------- one step
    // RN_min (reflective_newton.c:57-59)
    static double rn_min(double a, double b) { return a < b ? a : b; }
    // RN_test_sigma_l / _u (reflective_newton.c:460-470)
    double bound_term(double alpha, double gs, double shs, double sigma) const {
        return sigma * alpha * (gs + alpha * rn_min(shs, 0) / 2.);
    }

    void step_to(double alpha) {
        dev::copy(st_, n_, (const double*)Xo_, X_);
        axpy(X_, alpha, S_);
        dev::rn_reflect(st_, (int)idx_.size(), idxd_, lod_, upd_, X_);
    }

    bool touches_bound() {
        int np = dev::rn_touch(st_, (int)idx_.size(), idxd_, lod_, upd_, X_, part_);
        if (np == 0) return false;
        dev::finalize(st_, np, part_, scal_, 4, 0);
        return read(4) > 0;
    }

    // RN_iterate (reflective_newton.c:536-627)
    void iterate(bool on_dev) {
        const int n = n_;
        dev::copy(st_, n, (const double*)X_, Xo_);
        // M's products: on F's pattern (PCG / GMRES inner solvers, updated by
        // the step's device assembly), else M's own pattern, which keeps over
        // most Newton steps (values-only re-upload)
        if (!dMF_ && (!dM_ || !dM_->update_values(Mc_))) {
            dM_.reset();
            dM_.reset(new DevMatrix(Mc_));
        }
        DevMatrix& dM = dMF_ ? *dMF_ : *dM_;
        int status = subspace(on_dev, dM);
        // B = D g
        dev::copy(st_, n, (const double*)G_, Bv_);
        dev::mul(st_, n, Bv_, D_);
        const double nB = norm(Bv_);
        if (nB == 0) return;
        if (status == 1 && norm(V2_) < delta_) {
            dev::copy(st_, n, (const double*)V2_, S_);
        } else {
            normalize(V1_);
            double s = dot(V1_, V2_);
            axpy(V2_, -s, V1_);
            normalize(V2_);
            subspace_minimum(dM);
        }
        dev::mul(st_, n, S_, D_);
        // fixed quantities of the line search
        const double fo = functional(Xo_);
        const double gs = dot(G_, S_);
        const double shs = bilinear(*dH_, S_, S_);
        double alpha = 1.0;
        axpy(X_, alpha, S_);
        dev::rn_reflect(st_, (int)idx_.size(), idxd_, lod_, upd_, X_);
        if (!(functional(X_) - fo < bound_term(alpha, gs, shs, p_.sigma_l))) {
            stats_.line_searches++;
            const int max_iter = 10;
            int it = 0;
            double al = 0, ar = 1.;
            do {
                alpha = (al + ar) / 2.;
                step_to(alpha);
                double dqv = functional(X_) - fo;
                if (!(dqv < bound_term(alpha, gs, shs, p_.sigma_l))) {
                    ar = alpha;
                } else {
                    if (!(dqv > bound_term(alpha, gs, shs, p_.sigma_u)) && !(alpha > p_.rho))
                        al = alpha;
                    else
                        break;
                }
                it++;
            } while (it < max_iter);
            stats_.line_search_iter += it;
        }
        if (touches_bound()) {
            alpha -= p_.xi * nB;
            step_to(alpha);
        }
        const double d = functional(X_) - fo;
        const double pred = gs + rn_min(shs, 0);  // predicted_reduction (:472-474)
        const double contraction = d / pred;
        if (contraction > p_.trust_upper && d < 0) delta_ = rn_min(p_.radius_max, 2. * delta_);
        if (contraction < p_.trust_lower || d >= 0) delta_ = p_.radius_min > delta_ / 2. ? p_.radius_min : delta_ / 2.;
    }

    // subspace_minimum (reflective_newton.c:394-450) with b = B, basis V1, V2
    void subspace_minimum(DevMatrix& dM) {
        RootSearch rs;
        rs.bx = dot(V1_, Bv_);
        rs.by = dot(V2_, Bv_);
        rs.ax = bilinear(dM, V1_, V1_);
        rs.ay = bilinear(dM, V2_, V2_);
        rs.axy = bilinear(dM, V1_, V2_) + bilinear(dM, V2_, V1_);
        rs.delta = delta_;
        std::vector<double> roots = rs.roots();
        if (roots.empty()) throw NumericError("reflective Newton: no extremum on the trust-region circle");
        size_t m = 0;
        double mv = rs.q(roots[0]);
        for (size_t i = 1; i < roots.size(); i++)
            if (rs.q(roots[i]) < mv) {
                m = i;
                mv = rs.q(roots[i]);
            }
        dev::copy(st_, n_, (const double*)V1_, S_);
        dev::scale(st_, n_, S_, delta_ * std::cos(roots[m]));
        axpy(S_, delta_ * std::sin(roots[m]), V2_);
    }

    Csr H_;
    std::vector<double> b_, lo_, up_;
    std::vector<int> idx_;
    lina_rn_params p_;
    double delta_;
    uint64_t seed_;
    int n_ = 0;
    std::unique_ptr<DevMatrix> dH_;
    std::unique_ptr<DevMatrix> dM_;  // the scaled Hessian of the last Newton step
    std::unique_ptr<Solver> amg_;
    IluSystem ilu_;  // the ILU(0) inner solvers' system, refactored each Newton step
    std::unique_ptr<DevBuffers> B_;
    hipStream_t st_ = nullptr;
    double *X_ = nullptr, *Xo_ = nullptr, *G_ = nullptr, *D_ = nullptr, *J_ = nullptr, *V1_ = nullptr,
           *V2_ = nullptr, *bs_ = nullptr, *Bv_ = nullptr, *S_ = nullptr, *T_ = nullptr, *T2_ = nullptr;
    double *part_ = nullptr, *scal_ = nullptr, *lod_ = nullptr, *upd_ = nullptr, *bdev_ = nullptr;
    int* idxd_ = nullptr;
    double* host_ = nullptr;
    double last_curvature_ = 0;
    double t_m_ = 0, t_inner_ = 0, t_build_ = 0;
    int last_inner_ = 0;
    lina_rn_stats stats_{};
    // the scaled Hessian (PCG / GMRES inner solvers): F = H's pattern plus the
    // diagonal on the device; Mc_ the last host assembly (its values stale
    // after device steps until host_M()), mpos_ / mv_ its pattern's map and
    // device values
    Csr F_, Mc_;
    int *frp_ = nullptr, *fci_ = nullptr, *mflag_ = nullptr, *mpos_ = nullptr;
    double *fhv_ = nullptr, *mv_ = nullptr, *mvF_ = nullptr;
    std::unique_ptr<DevMatrix> dMF_;  // M on F's pattern (the products, line search, power method)
    uint8_t* fh_ = nullptr;
    std::unique_ptr<DevBuffers> mbufs_;
    bool mc_ready_ = false, mc_stale_ = false, mc_current_ = false;
    bool f_ilu_ = false, ilu_f_built_ = false;  // multicolour ILU on F's pattern
};

}  // namespace lina

// ------------------------------------------------------------------ C ABI

struct lina_rn {
    std::unique_ptr<lina::ReflectiveNewton> rn;
};

namespace {
template <typename F>
int rn_guard(F&& f) {
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

void lina_rn_params_default(lina_rn_params* p, int dof, int depth) {
    p->max_iter = 50;
    p->eps = 1e-8;
    p->sigma_l = 0.1;
    p->sigma_u = 0.9;
    p->rho = 0.1;
    p->delta = 250.;
    p->xi = 1.;
    p->trust_lower = 0.6;
    p->trust_upper = 1.1;
    p->radius_min = 1e-3;
    p->radius_max = 1000.;
    p->solver = LINA_RN_BICGSTAB;
    p->pcg_iter = 100;
    p->bi_eps = 1e-3;
    p->bi_max_iter = 10;
    p->gmres_m = 30;
    p->gmres_max_iter = 600;
    p->gmres_tol = 1e-10;
    lina_params_reference(&p->amg, dof, depth);
    p->amg.omega = 1.0;
    p->amg.amg_eps = 1e-4;
    p->amg.amg_max_iter = 20;
    p->seed = 0x5EED0007ULL;
    p->q_offset = 0.0;
    p->host_solver = nullptr;
    p->host_ctx = nullptr;
    p->ilu_ordering = LINA_ILU_NATURAL;
}

lina_rn* lina_rn_create(int n, const int* rowptr, const int* col, const double* val, const double* b, int nb,
                        const int* idx, const double* lower, const double* upper, const lina_rn_params* p) {
    lina_rn* h = nullptr;
    int rc = rn_guard([&]() {
        if (n <= 0 || !rowptr || !col || !val || !b || !p || nb < 0 || (nb > 0 && (!idx || !lower || !upper)))
            throw std::invalid_argument("lina_rn_create: bad arguments");
        if (p->solver < LINA_RN_BICGSTAB || p->solver > LINA_RN_HOST) throw std::invalid_argument("unknown solver");
        if (p->solver == LINA_RN_HOST && !p->host_solver) throw std::invalid_argument("LINA_RN_HOST needs host_solver");
        if (p->ilu_ordering != LINA_ILU_NATURAL && p->ilu_ordering != LINA_ILU_MULTICOLOUR)
            throw std::invalid_argument("unknown ILU ordering");
        if (p->solver == LINA_RN_GMRES && (p->gmres_m < 1 || p->gmres_max_iter < 1))
            throw std::invalid_argument("GMRES restart length and iteration limit must be positive");
        lina::Csr H;
        H.n = n;
        H.ncols = n;
        H.rp.assign(rowptr, rowptr + n + 1);
        if (H.rp[0] != 0) throw std::invalid_argument("rowptr[0] != 0");
        for (int i = 0; i < n; i++)
            if (H.rp[i + 1] <= H.rp[i]) throw std::invalid_argument("every row needs at least one entry");
        H.ci.assign(col, col + H.rp[n]);
        H.v.assign(val, val + H.rp[n]);
        for (int i = 0; i < n; i++)
            for (int q = H.rp[i]; q < H.rp[i + 1]; q++)
                if (H.ci[q] < 0 || H.ci[q] >= n || (q > H.rp[i] && H.ci[q] <= H.ci[q - 1]))
                    throw std::invalid_argument("columns out of range or not sorted");
        h = new lina_rn();
        h->rn.reset(new lina::ReflectiveNewton(H, b, nb, idx, lower, upper, *p));
        return 0;
    });
    if (rc < 0) {
        delete h;
        return nullptr;
    }
    return h;
}

void lina_rn_destroy(lina_rn* h) { delete h; }

int lina_rn_solve(lina_rn* h, double* x, lina_rn_stats* st) {
    return rn_guard([&]() {
        if (!h || !x) throw std::invalid_argument("null argument");
        return h->rn->solve(x, st);
    });
}

int lina_rn_scaled_hessian(int n, const int* rowptr, const int* col, const double* val, const double* D,
                           const double* G, const double* J, int* mrp, int* mci, double* mv) {
    return rn_guard([&]() {
        lina::Csr H;
        H.n = n;
        H.ncols = n;
        H.rp.assign(rowptr, rowptr + n + 1);
        H.ci.assign(col, col + H.rp[n]);
        H.v.assign(val, val + H.rp[n]);
        std::vector<double> d(D, D + n), g(G, G + n), j(J, J + n);
        lina::Csr M = lina::scaled_hessian(H, d, g, j);
        std::copy(M.rp.begin(), M.rp.end(), mrp);
        std::copy(M.ci.begin(), M.ci.end(), mci);
        std::copy(M.v.begin(), M.v.end(), mv);
        return M.nnz();
    });
}

int lina_rn_set_trust_region(lina_rn* h, double delta) {
    return rn_guard([&]() {
        if (!h) throw std::invalid_argument("null handle");
        h->rn->set_delta(delta);
        return 0;
    });
}

int lina_circle_minimum(double ax, double ay, double axy, double bx, double by, double delta, double* theta) {
    return rn_guard([&]() {
        if (!theta || !(delta > 0)) throw std::invalid_argument("lina_circle_minimum: bad arguments");
        lina::RootSearch rs;
        rs.ax = ax;
        rs.ay = ay;
        rs.axy = axy;
        rs.bx = bx;
        rs.by = by;
        rs.delta = delta;
        std::vector<double> roots = rs.roots();
        if (roots.empty()) throw lina::NumericError("no extremum on the trust-region circle");
        size_t m = 0;
        double mv = rs.q(roots[0]);
        for (size_t i = 1; i < roots.size(); i++)
            if (rs.q(roots[i]) < mv) {
                m = i;
                mv = rs.q(roots[i]);
            }
        *theta = roots[m];
        return (int)roots.size();
    });
}

}  // extern "C"
