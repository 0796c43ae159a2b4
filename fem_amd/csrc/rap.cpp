// Galerkin products of AMG_recompute_coarse_matrices on the device.
//
// AMG_recompute_coarse_matrices (Multigrid.c:1756-1780) forms, level by
// level, A_c = R (A P) with fixed P and R (get_restricted_matrix,
// Multigrid.c:475-484; sparse_product, linear_algebra.c:186-222).  In
// LinA's product a row of C = A B starts from sparse_zero's (i, 0.0), takes
// A(i,k) B(k,j) for the nonzero A(i,k) in row order and the nonzero B(k,j)
// in row order, adds each into (i, j) (the first one creating the entry),
// and ends with delete_zeros.
//
// While A's pattern holds, the products each entry of C receives, and their
// order, are fixed: a plan records them once per pattern, listing every
// structural entry of the changing factor (zero or not) and only the nonzero
// entries of the fixed one.  A listed product with a zero factor is an exact
// zero, which leaves any nonzero sum unchanged, so every entry that LinA
// keeps gets the value LinA computes, bit for bit; entries that sum to zero
// are the ones delete_zeros removes.  The device sums each entry's list in
// order (k_plan_sum_staged), first A P, then R (A P), and scatters the coarse
// entries into the level's CSR values, flagging any entry whose zero /
// nonzero state differs from the level's pattern (then the recompute takes
// the host path).
#include <hip/hip_runtime.h>
#include <omp.h>

#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "kernels.hpp"
#include "rap.hpp"
#include "solver.hpp"

namespace lina {

namespace {

// one row's share of a plan, built by one thread
struct RowPlan {
    std::vector<int> cols;    // entry columns, sorted
    std::vector<int> cnt;     // products per entry
    std::vector<uint8_t> seed;
    std::vector<int> ia, ib;  // products, grouped by entry, in LinA's order
};

// C = A B planned for row i: A's entries all (a_all) or only its nonzero
// ones, B's only its nonzero ones (b_all = false) or all; entries of C in
// columns < ncols; seed: the row starts with (i, 0.0) when i < ncols
void plan_row(int i, const Csr& A, bool a_all, const Csr& B, bool b_all, int ncols, std::vector<int>& slot,
              std::vector<int>& touched, RowPlan& out) {
    touched.clear();
    auto touch = [&](int c) {
        if (slot[c] < 0) {
            slot[c] = 0;
            touched.push_back(c);
        }
    };
    const bool seeded = i < ncols;
    if (seeded) touch(i);
    for (int p = A.rp[i]; p < A.rp[i + 1]; p++) {
        if (!a_all && A.v[p] == 0.0) continue;
        const int k = A.ci[p];
        for (int q = B.rp[k]; q < B.rp[k + 1]; q++) {
            if ((!b_all && B.v[q] == 0.0) || B.ci[q] >= ncols) continue;
            touch(B.ci[q]);
            slot[B.ci[q]]++;
        }
    }
    std::sort(touched.begin(), touched.end());
    const int m = (int)touched.size();
    size_t base = out.ia.size();
    std::vector<size_t> at(m);
    for (int e = 0; e < m; e++) {
        const int c = touched[e];
        out.cols.push_back(c);
        out.cnt.push_back(slot[c]);
        out.seed.push_back(seeded && c == i ? 1 : 0);
        at[e] = base;
        base += slot[c];
        slot[c] = -2 - e;  // entry index while filling
    }
    out.ia.resize(base);
    out.ib.resize(base);
    for (int p = A.rp[i]; p < A.rp[i + 1]; p++) {
        if (!a_all && A.v[p] == 0.0) continue;
        const int k = A.ci[p];
        for (int q = B.rp[k]; q < B.rp[k + 1]; q++) {
            if ((!b_all && B.v[q] == 0.0) || B.ci[q] >= ncols) continue;
            const int e = -2 - slot[B.ci[q]];
            out.ia[at[e]] = p;
            out.ib[at[e]] = q;
            at[e]++;
        }
    }
    for (int c : touched) slot[c] = -1;
}

}  // namespace

void plan_product(const Csr& A, bool a_all, const Csr& B, bool b_all, int ncols, ProductPlan& P) {
    const int n = A.n;
    const int nchunk = std::max(1, std::min(n, omp_get_max_threads() * 8));
    std::vector<RowPlan> parts(nchunk);
    std::vector<std::vector<int>> rowlen(nchunk);
#pragma omp parallel
    {
        std::vector<int> slot(std::max(ncols, 1), -1), touched;
#pragma omp for schedule(dynamic, 1)
        for (int c = 0; c < nchunk; c++) {
            const int r0 = (int)((long long)n * c / nchunk), r1 = (int)((long long)n * (c + 1) / nchunk);
            for (int i = r0; i < r1; i++) {
                const size_t before = parts[c].cols.size();
                plan_row(i, A, a_all, B, b_all, ncols, slot, touched, parts[c]);
                rowlen[c].push_back((int)(parts[c].cols.size() - before));
            }
        }
    }
    P.S.n = n;
    P.S.ncols = ncols;
    P.S.rp.assign(n + 1, 0);
    P.S.ci.clear();
    P.S.v.clear();
    P.ptr.assign(1, 0);
    P.seed.clear();
    P.ia.clear();
    P.ib.clear();
    int i = 0;
    for (int c = 0; c < nchunk; c++) {
        RowPlan& r = parts[c];
        for (int len : rowlen[c]) {
            P.S.rp[i + 1] = P.S.rp[i] + len;
            i++;
        }
        P.S.ci.insert(P.S.ci.end(), r.cols.begin(), r.cols.end());
        P.seed.insert(P.seed.end(), r.seed.begin(), r.seed.end());
        for (int k : r.cnt) P.ptr.push_back(P.ptr.back() + (int64_t)k);
        P.ia.insert(P.ia.end(), r.ia.begin(), r.ia.end());
        P.ib.insert(P.ib.end(), r.ib.begin(), r.ib.end());
        r = RowPlan();
    }
    P.S.v.assign(P.S.ci.size(), 0.0);
}

// the plan of level l (fine) -> l-1 (coarse) of H, with the coarse level's
// current pattern Ac for the scatter
void plan_galerkin(const HostLevel& fine, const Csr& Ac, RapLevelPlan& L) {
    plan_product(fine.A, true, fine.P, false, fine.P.ncols, L.ap);
    plan_product(fine.R, false, L.ap.S, true, Ac.n, L.ac);
    const int ne = (int)L.ac.S.ci.size();
    L.pos.assign(ne, -1);
    L.lone.assign(ne, 0);
    for (int I = 0; I < Ac.n; I++) {
        const int* b = Ac.ci.data() + Ac.rp[I];
        const int len = Ac.rp[I + 1] - Ac.rp[I];
        for (int e = L.ac.S.rp[I]; e < L.ac.S.rp[I + 1]; e++) {
            const int* f = std::lower_bound(b, b + len, L.ac.S.ci[e]);
            if (f != b + len && *f == L.ac.S.ci[e]) {
                L.pos[e] = Ac.rp[I] + (int)(f - b);
                L.lone[e] = len == 1 ? 1 : 0;
            }
        }
    }
}

size_t plan_bytes(const ProductPlan& P) {
    return P.ptr.size() * 4 + P.seed.size() + (P.ia.size() + P.ib.size()) * 4 + P.S.ci.size() * 8;
}

// ------------------------------------------------------------ Solver

// the device path needs the plain (non-batched) coarse inverse and no
// host-built lexicographic sweep packs; LINA_RAP=host keeps the host
// products (for comparison runs)
bool Solver::device_rap_ok() const {
    if (!uploaded() || cbatched_ || !seg_off_.empty() || !cpiv_ || nd_ < 0) return false;
    if ((size_t)nd_ * (size_t)nd_ > (size_t)INT32_MAX) return false;
    // the exact lexicographic SOR's packed records are rebuilt from host
    // values: such hierarchies keep the host path
    for (const DevLevel& d : L_)
        if (d.sor_packed) return false;
    const char* e = getenv("LINA_RAP");
    return !(e && std::strcmp(e, "host") == 0);
}

void Solver::build_rap() {
    const int depth = (int)H_.lv.size();
    hipStream_t st = stream_;
    std::unique_ptr<DeviceRap> R(new DeviceRap());
    DevBuffers& B = R->bufs;
    R->lv.assign(depth, DeviceRap::Lv());
    R->acmap.assign(depth, nullptr);
    for (int l = depth - 1; l >= 1; l--) {
        RapLevelPlan pl;
        plan_galerkin(H_.lv[l], H_.lv[l - 1].A, pl);
        DeviceRap::Lv& d = R->lv[l];
        d.nap = (int)pl.ap.S.ci.size();
        d.nac = (int)pl.ac.S.ci.size();
        d.ap_ptr = dupload(pl.ap.ptr, B, st);
        d.ap_a = dupload(pl.ap.ia, B, st);
        d.ap_p = dupload(pl.ap.ib, B, st);
        d.ap_seed = dupload(pl.ap.seed, B, st);
        d.ac_ptr = dupload(pl.ac.ptr, B, st);
        d.ac_r = dupload(pl.ac.ia, B, st);
        d.ac_ap = dupload(pl.ac.ib, B, st);
        d.ac_seed = dupload(pl.ac.seed, B, st);
        d.pos = dupload(pl.pos, B, st);
        d.lone = dupload(pl.lone, B, st);
        d.pv = dupload(H_.lv[l].P.v, B, st);
        d.rv = dupload(H_.lv[l].R.v, B, st);
        d.ap_v = dalloc<double>(d.nap, B);
        d.ac_v = dalloc<double>(d.nac, B);
        R->bytes += (double)plan_bytes(pl.ap) + (double)plan_bytes(pl.ac) + 5.0 * pl.pos.size();
        // the colour-permuted copy: its rows are A's in colour order
        const Csr& A = H_.lv[l].A;
        const std::vector<int>& col = colours_[l];
        const int nc = ncolours_[l];
        std::vector<int> cnt(nc + 1, 0), perm(A.n), map;
        for (int i = 0; i < A.n; i++) cnt[col[i] + 1]++;
        for (int c = 0; c < nc; c++) cnt[c + 1] += cnt[c];
        for (int i = 0; i < A.n; i++) perm[cnt[col[i]]++] = i;
        map.reserve(A.nnz());
        for (int r = 0; r < A.n; r++)
            for (int p = A.rp[perm[r]]; p < A.rp[perm[r] + 1]; p++) map.push_back(p);
        R->acmap[l] = dupload(map, B, st);
        hip_check(hipStreamSynchronize(st), "sync");  // the host plan dies here
    }
    // the coarsest level: dense index of each entry, each trivial row's diagonal
    const Csr& A0 = H_.lv[0].A;
    const CoarseLU& lu = H_.lu;
    std::vector<int> dmap(A0.n, -1), emap(A0.nnz(), -1), tmap(lu.trivial_idx.size(), -1);
    for (int g = 0; g < nd_; g++) dmap[lu.dense_idx[g]] = g;
    for (int i = 0; i < A0.n; i++)
        for (int p = A0.rp[i]; p < A0.rp[i + 1]; p++)
            if (dmap[i] >= 0 && dmap[A0.ci[p]] >= 0) emap[p] = dmap[i] * nd_ + dmap[A0.ci[p]];
    for (size_t t = 0; t < lu.trivial_idx.size(); t++) {
        const int i = lu.trivial_idx[t];
        for (int p = A0.rp[i]; p < A0.rp[i + 1]; p++)
            if (A0.ci[p] == i) tmap[t] = p;
    }
    R->dense_map = dupload(emap, B, st);
    R->tdiag_map = dupload(tmap, B, st);
    R->flag = dalloc<int>(1, B);
    hip_check(hipHostMalloc((void**)&R->host_flag, sizeof(int)), "hipHostMalloc");
    hip_check(hipStreamSynchronize(st), "sync");
    rap_ = std::move(R);
}

bool Solver::recompute_device(const Csr& A) {
    if (!rap_) build_rap();
    DevLevel& top = L_.back();
    if (A.nnz() != top.A.nnz) return false;
    hip_check(hipMemcpyAsync(top.A.v, A.v.data(), (size_t)A.nnz() * sizeof(double), hipMemcpyHostToDevice, stream_),
              "h2d");
    return recompute_device_core();
}

void Solver::recompute_device_values(const double* d_val) {
    if (!uploaded()) throw std::invalid_argument("recompute: the hierarchy is not on the device");
    DevLevel& top = L_.back();
    const size_t bytes = (size_t)top.A.nnz * sizeof(double);
    hipStream_t st = stream_;
    const bool timing = getenv("LINA_RN_TIMING") != nullptr;  // read per call: recomputes are rare
    if (device_rap_ok()) {
        const auto t0 = std::chrono::steady_clock::now();
        const bool built = rap_ != nullptr;
        if (!rap_) build_rap();
        hip_check(hipMemcpyAsync(top.A.v, d_val, bytes, hipMemcpyDeviceToDevice, st), "d2d");
        if (has_neg_signs_) {
            if (!neg_rows_) {
                std::vector<int> rows;
                for (int i = 0; i < n_; i++)
                    if (H_.signs[i] < 0) rows.push_back(i);
                nneg_ = (int)rows.size();
                neg_rows_ = dupload(rows, *bufs_, st);
            }
            dev::negate_rows(st, nneg_, neg_rows_, top.A.rp, top.A.v);
        }
        const bool ok = recompute_device_core();
        if (timing) {
            const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            fprintf(stderr, "[amg] recompute on the device from device values: %s %.3fs%s\n",
                    ok ? "done" : "pattern changed", sec, built ? "" : " (plan built)");
        }
        if (ok) {
            host_stale_ = true;
            host_top_stale_ = true;
            if (!refresh_isolated()) {
                sync_host_levels();
                upload();
            }
            return;
        }
    }
    // the host path from the caller's values
    std::vector<double> v(top.A.nnz);
    hip_check(hipMemcpyAsync(v.data(), d_val, bytes, hipMemcpyDeviceToHost, st), "d2h");
    hip_check(hipStreamSynchronize(st), "sync");
    recompute(v);
}

// the isolated rows' finest entries from the device into the host copy
void Solver::fetch_isolated_rows() {
    const Csr& A = H_.lv.back().A;
    if (!iso_pos_) {
        std::vector<int> pos;
        for (int i : H_.iso_rows)
            for (int p = A.rp[i]; p < A.rp[i + 1]; p++) pos.push_back(p);
        niso_pos_ = (int)pos.size();
        iso_pos_ = dupload(pos, *bufs_, stream_);
        iso_val_ = dalloc<double>(pos.size(), *bufs_);
        iso_pos_host_ = std::move(pos);
    }
    std::vector<double> v(niso_pos_);
    dev::gather<double>(stream_, niso_pos_, iso_pos_, L_.back().A.v, iso_val_);
    hip_check(hipMemcpyAsync(v.data(), iso_val_, (size_t)niso_pos_ * sizeof(double), hipMemcpyDeviceToHost, stream_),
              "d2h");
    hip_check(hipStreamSynchronize(stream_), "sync");
    Csr& Am = H_.lv.back().A;
    for (int k = 0; k < niso_pos_; k++) Am.v[iso_pos_host_[k]] = v[k];
}

bool Solver::recompute_device_core() {
    DeviceRap& R = *rap_;
    hipStream_t st = stream_;
    const int depth = (int)L_.size();
    dev::fill<int>(st, 1, R.flag, 0);
    for (int l = depth - 1; l >= 1; l--) {
        const DeviceRap::Lv& d = R.lv[l];
        dev::plan_sum(st, d.nap, d.ap_ptr, d.ap_a, d.ap_p, d.ap_seed, L_[l].A.v, d.pv, d.ap_v);
        dev::plan_sum(st, d.nac, d.ac_ptr, d.ac_r, d.ac_ap, d.ac_seed, d.rv, d.ap_v, d.ac_v);
        dev::plan_scatter(st, d.nac, d.ac_v, d.pos, d.lone, L_[l - 1].A.v, R.flag);
    }
    hip_check(hipMemcpyAsync(R.host_flag, R.flag, sizeof(int), hipMemcpyDeviceToHost, st), "d2h");
    hip_check(hipStreamSynchronize(st), "sync");
    if (*R.host_flag) return false;
    // what each level derives from its values: the FP32 copy, the rows' l1
    // norms, the colour-permuted copies
    for (int l = 0; l < depth; l++) {
        DevLevel& d = L_[l];
        dev::convert<double, float>(st, d.A.nnz, d.A.v, d.Af.v);
        dev::row_l1<double>(st, d.n, d.A.rp, d.A.v, d.A.l1);
        dev::row_l1<float>(st, d.n, d.A.rp, d.A.v, d.Af.l1);
        if (l == 0) continue;
        dev::gather<double>(st, d.A.nnz, R.acmap[l], d.A.v, d.Ac.v);
        dev::convert<double, float>(st, d.A.nnz, d.Ac.v, d.Acf.v);
        dev::gather<double>(st, d.n, d.perm, d.A.l1, d.Ac.l1);
        dev::gather<float>(st, d.n, d.perm, d.Af.l1, d.Acf.l1);
    }
    // the coarsest direct solve: the dense rows' matrix, inverted in place
    dev::fill<double>(st, nd_ * nd_, inv_, 0.0);
    dev::scatter(st, L_[0].A.nnz, R.dense_map, L_[0].A.v, inv_);
    dev::gather<double>(st, nt_, R.tdiag_map, L_[0].A.v, tdiag_);
    dev::gauss_jordan_inverse(st, nd_, inv_, cpiv_);
    if (ctiles_) dev::inverse_to_tiles<double>(st, nd_, ntiles_, cti_, ctj_, inv_, ctiles_);
    hip_check(hipStreamSynchronize(st), "recompute sync");
    return true;
}

// the device's coarse values back into the host hierarchy (and the coarse
// dense system rebuilt from them)
void Solver::sync_host_levels() {
    if (host_top_stale_) {
        Csr& A = H_.lv.back().A;
        hip_check(hipMemcpyAsync(A.v.data(), L_.back().A.v, (size_t)A.nnz() * sizeof(double), hipMemcpyDeviceToHost,
                                 stream_), "d2h");
        hip_check(hipStreamSynchronize(stream_), "sync");
        host_top_stale_ = false;
    }
    if (!host_stale_) return;
    const int depth = (int)H_.lv.size();
    for (int l = 0; l < depth - 1; l++) {
        Csr& A = H_.lv[l].A;
        hip_check(hipMemcpyAsync(A.v.data(), L_[l].A.v, (size_t)A.nnz() * sizeof(double), hipMemcpyDeviceToHost,
                                 stream_), "d2h");
    }
    hip_check(hipStreamSynchronize(stream_), "sync");
    coarse_lu(H_.lv[0].A, H_.lu);
    host_stale_ = false;
}

const HostHierarchy& Solver::host_synced() {
    sync_host_levels();
    return H_;
}

}  // namespace lina
