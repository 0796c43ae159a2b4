// Krylov drivers beside AMG-PCG: ILU(0)-preconditioned pcg and GMRES(m),
// and BiCGStab with the AMG preconditioner.  Vectors stay on the device;
// the host reads the scalars that steer the iteration.
//
//   pcg with PCG_preconlr      PCGS.c:61-108, 163-233
//   GMRES(m), right ILU(0)     the restarted GMRES that FEM.c drives
//                              (FEM.c:730-745, CheckOnRestart): classical
//                              Gram-Schmidt, Givens QR of the Hessenberg
//                              matrix, convergence tested at restarts; own
//                              implementation after Saad, Iterative Methods
//                              for Sparse Linear Systems, Alg. 6.9 / 9.5
//   BiCGStab_solve             BiCGStab.c:115-198

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstdio>
#include <chrono>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <vector>

#include "ilu.hpp"
#include "kernels.hpp"
#include "solver.hpp"

namespace lina {

template <typename T>
static T* kalloc(size_t n, DevBuffers& owner) {
    if (n == 0) n = 1;
    void* p = nullptr;
    hip_check(hipMalloc(&p, n * sizeof(T)), "hipMalloc");
    owner.ptrs.push_back(p);
    return (T*)p;
}

template <typename T>
static T* kupload(const std::vector<T>& h, DevBuffers& owner, hipStream_t st) {
    T* d = kalloc<T>(h.size(), owner);
    if (!h.empty()) hip_check(hipMemcpyAsync(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, st), "upload");
    return d;
}

// dependency levels of a triangular sweep over nonzero off-diagonal values
// level schedule of a triangular factor: level(i) = 1 + max level(j) over
// its off-diagonal entries (entries holding 0.0 skipped unless
// pattern_only, for factors whose values are not known yet)
static void tri_levels(const Csr& T, bool upper, std::vector<int>& off, std::vector<int>& rows,
                       bool pattern_only = false) {
    const int n = T.n;
    std::vector<int> lev(n, 0);
    int nl = 0;
    for (int t = 0; t < n; t++) {
        int i = upper ? n - 1 - t : t;
        int l = 0;
        for (int p = T.rp[i]; p < T.rp[i + 1]; p++) {
            int j = T.ci[p];
            if ((!pattern_only && T.v[p] == 0.0) || j == i) continue;
            if ((!upper && j < i) || (upper && j > i)) l = std::max(l, lev[j] + 1);
        }
        lev[i] = l;
        nl = std::max(nl, l + 1);
    }
    off.assign(nl + 1, 0);
    for (int i = 0; i < n; i++) off[lev[i] + 1]++;
    for (int l = 0; l < nl; l++) off[l + 1] += off[l];
    rows.assign(n, 0);
    std::vector<int> fill(off.begin(), off.end() - 1);
    for (int i = 0; i < n; i++) rows[fill[lev[i]]++] = i;
}

// The entries of a triangular factor's row that its level schedule counted
// (all but the diagonal -- the last entry of an L row, the first of a U row --
// and, unless pattern_only, the zero-valued ones); -1 for a factor with an
// entry on the wrong side of the diagonal or an empty row, else the largest
// count of a row.
static int tri_max_count(const Csr& T, bool upper, bool pattern_only) {
    int maxc = 0;
    for (int i = 0; i < T.n; i++) {
        if (T.rp[i + 1] <= T.rp[i]) return -1;
        const int d = upper ? T.rp[i] : T.rp[i + 1] - 1;
        int c = 0;
        for (int p = T.rp[i]; p < T.rp[i + 1]; p++) {
            if (p == d) continue;
            const int j = T.ci[p];
            if (upper ? j <= i : j >= i) return -1;
            c += pattern_only || T.v[p] != 0.0;
        }
        maxc = std::max(maxc, c);
    }
    return maxc;
}

// A triangular factor packed by level for dev::tri_solve_pair (records of
// MAXP value slots): rows in the level schedule's order (off, rows), the
// counted entries in row order, which all lie in earlier levels.  A
// dependency whose result is still in the LDS ring when its reader's level
// ends (fewer than TRI_RING phase positions back) is read there, any other
// from outpos in global memory.  The values are filled by pack_values.
template <int MAXP>
static bool tri_pack(const Csr& T, bool upper, bool pattern_only, const std::vector<int>& off,
                     const std::vector<int>& rows, const std::vector<int>& outpos, DevBuffers& B, hipStream_t st,
                     IluSystem::TriDev& D) {
    const int n = T.n, nlv = (int)off.size() - 1;
    std::vector<int> pos(n), lev(n);
    for (int L = 0; L < nlv; L++)
        for (int t = off[L]; t < off[L + 1]; t++) {
            pos[rows[t]] = t;
            lev[rows[t]] = L;
        }
    const size_t N = (size_t)n;
    std::vector<int> full(std::max(nlv, 1), 0), dmap(n), vmap(MAXP * N, -1), ovp(n, 0), ovmap, ovsrc;
    std::vector<dev::TriRec<MAXP>> rec(n);
    for (int t = 0; t < n; t++) {
        const int i = rows[t], L = lev[i], e = off[L + 1];
        const int d = upper ? T.rp[i] : T.rp[i + 1] - 1;
        dev::TriRec<MAXP>& r = rec[t];
        r.diag = 0.0;
        for (int k = 0; k < MAXP; k++) {
            r.val[k] = 0.0;
            r.src[k] = dev::TRI_RING;  // the ring's zero slot
        }
        r.orow = outpos[i];
        dmap[t] = d;
        ovp[t] = (int)ovmap.size();
        int k = 0;
        for (int p = T.rp[i]; p < T.rp[i + 1]; p++) {
            if (p == d || !(pattern_only || T.v[p] != 0.0)) continue;
            const int j = T.ci[p];
            int sj;
            if (e - pos[j] <= dev::TRI_RING) {
                sj = pos[j] & (dev::TRI_RING - 1);
            } else {
                if (MAXP == 3) return false;  // 16-bit sources hold ring slots only
                sj = -(outpos[j] + 1);
                full[L] = 1;
            }
            if (k < MAXP) {
                vmap[k * N + t] = p;
                r.src[k] = sj;
            } else {
                ovmap.push_back(p);
                ovsrc.push_back(sj);
            }
            k++;
        }
        r.cnt = k;
    }
    std::vector<int> offf(off);
    for (int L = 0; L < nlv; L++)
        if (full[L]) offf[L] |= (int)0x80000000u;
    dev::TriPack& P = D.p;
    P.n = n;
    P.nlv = nlv;
    P.maxp = MAXP;
    P.off = kupload(offf, B, st);
    D.rec = kupload(rec, B, st);
    P.rec = D.rec;
    P.ovp = kupload(ovp, B, st);
    P.ovsrc = kupload(ovsrc, B, st);
    D.nov = (int)ovmap.size();
    D.vmap = kupload(vmap, B, st);
    D.dmap = kupload(dmap, B, st);
    D.ovmap = kupload(ovmap, B, st);
    D.ovval = kalloc<double>(ovmap.size(), B);
    P.ovval = D.ovval;
    return true;
}

// blocks: also the LDS row blocks of the SpMV (the factors need only the CSR)
static dev::DMat<double> upload_csr(const Csr& A, DevBuffers& B, hipStream_t st, bool blocks = true) {
    dev::DMat<double> D;
    D.n = A.n;
    D.ncols = A.ncols;
    D.nnz = A.nnz();
    D.rp = kupload(A.rp, B, st);
    D.ci = kupload(A.ci, B, st);
    D.v = kupload(A.v, B, st);
    if (!blocks) return D;
    // LDS row blocks for the SpMV (same builder rule as the AMG levels)
    std::vector<int> starts;
    build_blocks(A, 0, A.n, dev::block_nnz<double>(), starts);
    starts.push_back(A.n);
    if (A.n == 0) starts.assign(1, 0);
    D.nblk = (int)starts.size() - 1;
    D.bk = kupload(block_pairs(A, starts), B, st);
    D.lpr = dev::lanes_per_row(A.n > 0 ? (double)A.nnz() / A.n : 0.0);
    return D;
}

// ------------------------------------------------------------- IluSystem

IluSystem::IluSystem() {}

IluSystem::~IluSystem() {
    if (stream_) (void)hipStreamSynchronize(stream_);
    if (host_) (void)hipHostFree(host_);
    bufs_.reset();
    if (stream_) (void)hipStreamDestroy(stream_);
}

void IluSystem::set_ordering(int ordering) {
    if (ordering != 0 && ordering != 1) throw std::invalid_argument("unknown ILU ordering");
    if (ordering != ordering_) {
        ordering_ = ordering;
        device_ready_ = false;
        opat_ = Csr();
        ppat_ = Csr();
        perm_.clear();
        vpos_.clear();
    }
}

// P A P^T for the multicolour order: P from the greedy colouring of A's
// pattern, explicit zeros included (rows colour by colour, natural order
// within a colour), each permuted row's entries sorted by their new column
Csr IluSystem::permuted(const Csr& A) {
    const bool same = A.n == opat_.n && A.rp == opat_.rp && A.ci == opat_.ci;
    if (!same) {
        const int n = A.n;
        int nc = 0;
        // colour the pattern, explicit zeros included: ILU(0) may fill them,
        // and then rows of one colour would depend on each other
        Csr pat;
        pat.n = pat.ncols = n;
        pat.rp = A.rp;
        pat.ci = A.ci;
        pat.v.assign(A.ci.size(), 1.0);
        const std::vector<int> col = greedy_colouring(pat, &nc);
        std::vector<int> cnt(nc + 1, 0), iperm(n);
        for (int i = 0; i < n; i++) cnt[col[i] + 1]++;
        for (int c = 0; c < nc; c++) cnt[c + 1] += cnt[c];
        perm_.assign(n, 0);
        for (int i = 0; i < n; i++) perm_[cnt[col[i]]++] = i;
        for (int r = 0; r < n; r++) iperm[perm_[r]] = r;
        ppat_ = Csr();
        ppat_.n = ppat_.ncols = n;
        ppat_.rp.assign(n + 1, 0);
        ppat_.ci.resize(A.nnz());
        vpos_.resize(A.nnz());
        std::vector<std::pair<int, int>> row;
        for (int r = 0; r < n; r++) {
            const int i = perm_[r];
            row.clear();
            for (int p = A.rp[i]; p < A.rp[i + 1]; p++) row.emplace_back(iperm[A.ci[p]], p);
            std::sort(row.begin(), row.end());
            int k = ppat_.rp[r];
            for (const auto& e : row) {
                ppat_.ci[k] = e.first;
                vpos_[k++] = e.second;
            }
            ppat_.rp[r + 1] = k;
        }
        opat_ = Csr();
        opat_.n = opat_.ncols = n;
        opat_.rp = A.rp;
        opat_.ci = A.ci;
        device_ready_ = false;
    }
    Csr P = ppat_;
    P.v.resize(vpos_.size());
    for (size_t k = 0; k < vpos_.size(); k++) P.v[k] = A.v[vpos_[k]];
    return P;
}

void IluSystem::build(const Csr& A, bool upload) {
    if (ordering_) build_p(permuted(A), upload);
    else build_p(A, upload);
}

void IluSystem::build_p(const Csr& A, bool upload) {
    A_ = A;
    device_ready_ = false;
    host_factors_ = true;
    Csr L, U;
    ilu_ok_ = ilu0(A, L, U);
    L_ = std::move(L);
    U_ = std::move(U);
    n_ = A.n;
    if (upload) this->upload();
}

// the patterns ilu0 produces for A (rows strictly ascending): L row i = the
// lower columns then i, U row i = i then the upper columns; false when a row
// is not sorted
static bool ilu_patterns(const Csr& A, Csr& L, Csr& U) {
    const int n = A.n;
    L = Csr();
    U = Csr();
    L.n = L.ncols = U.n = U.ncols = n;
    L.rp.assign(n + 1, 0);
    U.rp.assign(n + 1, 0);
    for (int i = 0; i < n; i++) {
        for (int p = A.rp[i]; p < A.rp[i + 1]; p++) {
            const int j = A.ci[p];
            if (p > A.rp[i] && j <= A.ci[p - 1]) return false;
            if (j < i) L.ci.push_back(j);
        }
        L.ci.push_back(i);
        L.rp[i + 1] = (int)L.ci.size();
        U.ci.push_back(i);
        for (int p = A.rp[i]; p < A.rp[i + 1]; p++)
            if (A.ci[p] > i) U.ci.push_back(A.ci[p]);
        U.rp[i + 1] = (int)U.ci.size();
    }
    L.v.assign(L.ci.size(), 0.0);
    U.v.assign(U.ci.size(), 0.0);
    return true;
}

void IluSystem::refactor(const Csr& A) {
    if (ordering_) refactor_p(permuted(A));
    else refactor_p(A);
}

void IluSystem::refactor_p(const Csr& A) {
    const bool same = device_ready_ && A.n == A_.n && A.rp == A_.rp && A.ci == A_.ci;
    if (!same) {
        Csr L, U;
        if (!ilu_patterns(A, L, U)) {
            device_ready_ = false;
            build_p(A, true);
            return;
        }
        const char* tenv = getenv("LINA_RN_TIMING");
        const bool timing = tenv && atoi(tenv) != 0;
        const auto t0 = std::chrono::steady_clock::now();
        A_ = A;
        L_ = std::move(L);
        U_ = std::move(U);
        n_ = A.n;
        upload(true);
        const auto t1 = std::chrono::steady_clock::now();
        plan_ = build_plan();
        device_ready_ = true;
        if (timing) {
            auto sec = [](auto a, auto b) { return std::chrono::duration<double>(b - a).count(); };
            fprintf(stderr, "[ilu] new pattern: upload %.3fs, plan %.3fs\n", sec(t0, t1),
                    sec(t1, std::chrono::steady_clock::now()));
        }
    } else {
        // A_ keeps the pattern; its values are not refreshed (only dA_'s are)
        hip_check(hipMemcpyAsync(dA_.v, A.v.data(), A.v.size() * sizeof(double), hipMemcpyHostToDevice, stream_),
                  "h2d");
    }
    numeric();
}

bool IluSystem::refactor_device(const double* d_val) {
    if (!device_ready_ || !stream_) return false;
    const int nnz = A_.nnz();
    if (dvpos_) dev::permute_idx(stream_, nnz, dvpos_, d_val, dA_.v);
    else hip_check(hipMemcpyAsync(dA_.v, d_val, (size_t)nnz * sizeof(double), hipMemcpyDeviceToDevice, stream_), "d2d");
    numeric();
    return true;
}

// the numeric factorisation of dA_'s values on the symbolic structure in place
void IluSystem::numeric() {
    host_[0] = 0.0;
    hip_check(hipMemcpyAsync(fail_, host_, sizeof(int), hipMemcpyHostToDevice, stream_), "h2d");
    if (plan_ && wide_)
        dev::ilu0_planned_levels(stream_, hoff_l_, pm1_, pm2_, pmask_, ppu_, dA_.v, dL_.v, dU_.v, fail_);
    else if (plan_)
        dev::ilu0_planned(stream_, nlv_l_, lvl_off_, pm1_, pm2_, pmask_, ppu_, dA_.v, dL_.v, dU_.v, fail_);
    else
        dev::ilu0_numeric(stream_, nlv_l_, lvl_off_, lvl_rows_, dA_.rp, dA_.ci, dA_.v, dL_.rp, dL_.ci, dL_.v,
                          dU_.rp, dU_.ci, dU_.v, fail_);
    pack_values();
    int fail = 0;
    hip_check(hipMemcpyAsync(&host_[1], fail_, sizeof(int), hipMemcpyDeviceToHost, stream_), "d2h");
    hip_check(hipStreamSynchronize(stream_), "sync");
    std::memcpy(&fail, &host_[1], sizeof(int));
    ilu_ok_ = fail ? 0 : 1;
    host_factors_ = false;
}

// the symbolic plan of dev::ilu0_planned for A_ and the patterns L_, U_ (rows
// sorted; L's level schedule hoff_l_/hrows_l_): false when a row has more
// entries than the kernel's register slots (the searching kernel runs then)
bool IluSystem::build_plan() {
    const int n = n_;
    std::vector<int4> m1(n), m2(n);
    std::vector<int> mask(A_.nnz(), 0), npair(n + 1, 0);
    bool fits = true;
    // U row k: the diagonal first, then the upper columns ascending; the
    // entry of column j, or -1
    auto u_entry = [&](int k, int j) {
        const int* b = U_.ci.data() + U_.rp[k] + 1;
        const int* e = U_.ci.data() + U_.rp[k + 1];
        const int* f = std::lower_bound(b, e, j);
        return (f != e && *f == j) ? (int)(f - U_.ci.data()) : -1;
    };
    // pass 1 (rows independent): counts, masks, row descriptors
#pragma omp parallel for schedule(dynamic, 4096) reduction(&& : fits)
    for (int t = 0; t < n; t++) {
        const int i = hrows_l_[t];
        const int a0 = A_.rp[i], na = A_.rp[i + 1] - a0;
        int nl = 0, hasd = 0;
        for (int m = 0; m < na; m++) {
            nl += A_.ci[a0 + m] < i;
            hasd |= A_.ci[a0 + m] == i;
        }
        if (na > dev::ILU_MAXA || nl > dev::ILU_MAXL) {
            fits = false;
            continue;
        }
        m1[t] = make_int4(a0, na, L_.rp[i], U_.rp[i]);
        int cnt = 0;
        for (int m = 0; m < na; m++) {
            const int j = A_.ci[a0 + m], lim = std::min(m, nl);
            int mk = 0;
            for (int q = 0; q < lim; q++)
                if (u_entry(A_.ci[a0 + q], j) >= 0) {
                    mk |= 1 << q;
                    cnt++;
                }
            mask[a0 + m] = mk;
        }
        m2[t] = make_int4(nl, hasd, 0, 0);
        npair[t + 1] = cnt;
    }
    if (!fits) return false;
    for (int t = 0; t < n; t++) npair[t + 1] += npair[t];
    std::vector<int> pu(npair[n]);
    // pass 2: the pairs' U entries, slot order, entry after entry
#pragma omp parallel for schedule(dynamic, 4096)
    for (int t = 0; t < n; t++) {
        const int a0 = m1[t].x, na = m1[t].y, nl = m2[t].x;
        int r = npair[t];
        m2[t].z = r;
        for (int m = 0; m < na; m++) {
            const int mk = mask[a0 + m], j = A_.ci[a0 + m];
            for (int q = 0; q < std::min(m, nl); q++)
                if ((mk >> q) & 1) pu[r++] = u_entry(A_.ci[a0 + q], j);
        }
    }
    DevBuffers& B = *bufs_;
    pm1_ = kupload(m1, B, stream_);
    pm2_ = kupload(m2, B, stream_);
    pmask_ = kupload(mask, B, stream_);
    ppu_ = kupload(pu, B, stream_);
    hip_check(hipStreamSynchronize(stream_), "plan upload");
    return true;
}

void IluSystem::sync_factors() {
    if (host_factors_ || !stream_) return;
    hip_check(hipMemcpyAsync(L_.v.data(), dL_.v, L_.v.size() * sizeof(double), hipMemcpyDeviceToHost, stream_),
              "d2h");
    hip_check(hipMemcpyAsync(U_.v.data(), dU_.v, U_.v.size() * sizeof(double), hipMemcpyDeviceToHost, stream_),
              "d2h");
    hip_check(hipStreamSynchronize(stream_), "sync");
    host_factors_ = true;
}

void IluSystem::set_factors(const Csr& L, const Csr& U, bool up) {
    if (L.n != U.n) throw std::invalid_argument("L and U sizes differ");
    set_ordering(0);  // factors handed over are in the caller's order
    A_ = Csr();
    device_ready_ = false;
    host_factors_ = true;
    L_ = L;
    U_ = U;
    ilu_ok_ = 1;
    n_ = L.n;
    if (up) upload();
}

// pattern_levels: level schedules from the factors' patterns (their values
// are computed on the device afterwards)
void IluSystem::upload(bool pattern_levels_in) {
    // the multicolour order's wide sweeps read every entry of a row, zeros
    // included, so their levels come from the patterns (a level schedule
    // from the values lets a row read an entry that a zero factor value
    // ties to a later row)
    const bool pattern_levels = pattern_levels_in || ordering_ != 0;
    if (!stream_) {
        hip_check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate");
        hip_check(hipHostMalloc((void**)&host_, 64 * sizeof(double)), "hipHostMalloc");
    }
    // a re-upload frees every earlier buffer, the GMRES basis allocated on
    // first use included: nothing may still run on them, and no pointer to
    // them may survive
    hip_check(hipStreamSynchronize(stream_), "sync");
    bufs_.reset(new DevBuffers());
    vbasis_ = nullptr;
    hcols_ = nullptr;
    mb_ = 0;
    DevBuffers& B = *bufs_;
    hipStream_t st = stream_;
    if (A_.n > 0) dA_ = upload_csr(A_, B, st);
    dL_ = upload_csr(L_, B, st, false);
    dU_ = upload_csr(U_, B, st, false);
    std::vector<int> off, rows;
    std::vector<int> offu, rowsu;
    tri_levels(L_, false, off, rows, pattern_levels);
    hoff_l_ = off;
    hrows_l_ = rows;
    plan_ = false;
    nlv_l_ = (int)off.size() - 1;
    lvl_off_ = kupload(off, B, st);
    lvl_rows_ = kupload(rows, B, st);
    tri_levels(U_, true, offu, rowsu, pattern_levels);
    nlv_u_ = (int)offu.size() - 1;
    lvu_off_ = kupload(offu, B, st);
    lvu_rows_ = kupload(rowsu, B, st);
    hoff_u_ = offu;
    const int n = n_;
    for (double** p : {&vX_, &vB_, &vR_, &vP_, &vQ_, &vZ_, &vW_, &vY_, &vT_}) *p = kalloc<double>(n, B);
    dperm_ = nullptr;
    dvpos_ = nullptr;
    vXo_ = vBo_ = vPz_ = vPw_ = nullptr;
    if (ordering_ && (int)perm_.size() == n) {
        dperm_ = kupload(perm_, B, st);
        if ((int)vpos_.size() == A_.nnz() && A_.nnz() > 0) dvpos_ = kupload(vpos_, B, st);
        for (double** p : {&vXo_, &vBo_, &vPz_, &vPw_}) *p = kalloc<double>(n, B);
    }
    // a few wide levels (the multicolour order's): one launch per level on
    // the whole GPU instead of the one-workgroup level packs
    wide_ = pattern_levels && n >= 4096 && nlv_l_ <= 64 && nlv_u_ <= 64;
    // level packs: the forward sweep writes y in the backward sweep's order
    std::vector<int> posu(n), natural(n);
    for (int t = 0; t < n; t++) posu[rowsu[t]] = t;
    for (int i = 0; i < n; i++) natural[i] = i;
    tf_ = TriDev();
    tb_ = TriDev();
    const int cl = tri_max_count(L_, false, pattern_levels), cu = tri_max_count(U_, true, pattern_levels);
    packed_ = n > 0 && cl >= 0 && cu >= 0 && !wide_;
    if (packed_) {
        // one record width for both sweeps (one kernel runs them); 48-byte
        // records when every row fits 3 entries whose sources are in the ring
        bool done = false;
        if (std::max(cl, cu) <= 3)
            done = tri_pack<3>(L_, false, pattern_levels, off, rows, posu, B, st, tf_) &&
                   tri_pack<3>(U_, true, pattern_levels, offu, rowsu, natural, B, st, tb_);
        if (done) {
        } else if (std::max(cl, cu) <= 4) {
            tri_pack<4>(L_, false, pattern_levels, off, rows, posu, B, st, tf_);
            tri_pack<4>(U_, true, pattern_levels, offu, rowsu, natural, B, st, tb_);
        } else {
            tri_pack<8>(L_, false, pattern_levels, off, rows, posu, B, st, tf_);
            tri_pack<8>(U_, true, pattern_levels, offu, rowsu, natural, B, st, tb_);
        }
        permL_ = lvl_rows_;
        vZp_ = kalloc<double>(n, B);
        vYu_ = kalloc<double>(n, B);
        pack_values();
    }
    part_ = kalloc<double>(std::max(4096, dA_.nblk + 16) * 40, B);
    scal_ = kalloc<double>(64, B);
    fail_ = kalloc<int>(1, B);
    dev::fill(st, 64, scal_, 0.0);
    hip_check(hipStreamSynchronize(st), "upload sync");
}

void IluSystem::pack_values() {
    if (!packed_) return;
    dev::tri_pack_values(stream_, n_, tf_.p.maxp, tf_.vmap, tf_.dmap, tf_.nov, tf_.ovmap, dL_.v, tf_.rec, tf_.ovval);
    dev::tri_pack_values(stream_, n_, tb_.p.maxp, tb_.vmap, tb_.dmap, tb_.nov, tb_.ovmap, dU_.v, tb_.rec, tb_.ovval);
}

void IluSystem::apply(const double* z, double* w) {
    if (!dperm_) {
        apply_p(z, w);
        return;
    }
    dev::permute(stream_, n_, dperm_, z, vPz_, 0);
    apply_p(vPz_, vPw_);
    dev::permute(stream_, n_, dperm_, vPw_, w, 1);
}

void IluSystem::apply_p(const double* z, double* w) {
    // PCG_preconlr: y = L^-1 z, w = U^-1 y
    if (packed_) {
        dev::tri_solve_pair(stream_, tf_.p, tb_.p, permL_, z, vZp_, vYu_, w);
        return;
    }
    if (wide_) {
        dev::tri_levels_wide(stream_, dL_.rp, dL_.ci, dL_.v, hoff_l_, lvl_rows_, z, vT_, 0);
        dev::tri_levels_wide(stream_, dU_.rp, dU_.ci, dU_.v, hoff_u_, lvu_rows_, vT_, w, 1);
        return;
    }
    dev::trsv(stream_, dL_.rp, dL_.ci, dL_.v, nlv_l_, lvl_off_, lvl_rows_, z, vT_, 0);
    dev::trsv(stream_, dU_.rp, dU_.ci, dU_.v, nlv_u_, lvu_off_, lvu_rows_, vT_, w, 1);
}

double IluSystem::dot(const double* a, const double* b) {
    int np = dev::dot_part<double>(stream_, n_, a, b, part_);
    dev::finalize(stream_, np, part_, scal_, 0, 0);
    hip_check(hipMemcpyAsync(host_, scal_, sizeof(double), hipMemcpyDeviceToHost, stream_), "d2h");
    hip_check(hipStreamSynchronize(stream_), "sync");
    return host_[0];
}

void IluSystem::matvec(const double* x, double* y) {
    dev::spmv<double, double>(stream_, dA_, x, nullptr, y, dev::MODE_SET, nullptr, nullptr);
}

int IluSystem::pcg_dev(double* X, double* B, double tol, int max_iter, int stop, double* residuum) {
    if (!dperm_) return pcg_core(X, B, tol, max_iter, stop, residuum);
    if (!ilu_ok_) throw NumericError("ILU(0): zero pivot");
    dev::permute(stream_, n_, dperm_, X, vXo_, 0);
    dev::permute(stream_, n_, dperm_, B, vBo_, 0);
    const int it = pcg_core(vXo_, vBo_, tol, max_iter, stop, residuum);
    dev::permute(stream_, n_, dperm_, vXo_, X, 1);
    dev::permute(stream_, n_, dperm_, vBo_, B, 1);
    hip_check(hipStreamSynchronize(stream_), "sync");
    return it;
}

// pcg (PCGS.c:163-233) with precon = PCG_preconlr and mat_mult = PCG_mult
int IluSystem::pcg_core(double* X, double* B, double tol, int max_iter, int stop, double* residuum) {
    if (!ilu_ok_) throw NumericError("ILU(0): zero pivot");
    hipStream_t st = stream_;
    const int n = n_;
    double h = std::sqrt(dot(B, B));
    if (h == 0) h = 1;
    const double inv = (double)1 / h;
    dev::axpby(st, n, inv, B, 0.0, B);
    dev::axpby(st, n, inv, X, 0.0, X);
    double *R = vR_, *P = vP_, *Q = vQ_, *Z = vZ_, *W = vW_;
    dev::spmv<double, double>(st, dA_, X, B, R, dev::MODE_RES, nullptr, nullptr);
    apply_p(R, W);
    double r0 = std::sqrt(dot(W, W));
    if (r0 == 0) return 0;
    double r = r0, s = 0, s_old = 0;
    int i = 0;
    do {
        apply_p(R, Z);
        s_old = s;
        s = dot(R, Z);
        if (i > 0) {
            double b = s / s_old;
            dev::axpby(st, n, b, P, 1.0, Z);  // P*b + 1*Z, rounded as scalar_mult then vector_add
        } else {
            dev::copy(st, n, (const double*)Z, P);
        }
        matvec(P, Q);
        double a = s / dot(P, Q);
        dev::axpby(st, n, 1.0, X, a, P);
        dev::axpby(st, n, 1.0, R, -a, Q);
        r = std::sqrt(dot(R, R));
        i++;
    } while ((stop == LINA_STOP_RELRES ? r : r / r0) > tol && i < max_iter);
    apply_p(R, W);
    double rw = std::sqrt(dot(W, W));
    dev::axpby(st, n, h, B, 0.0, B);
    dev::axpby(st, n, h, X, 0.0, X);
    hip_check(hipStreamSynchronize(st), "sync");
    if (residuum) *residuum = rw / r0;
    if (std::isnan(rw / r0)) throw NumericError("pcg: NaN residual");
    return i;
}

// Givens QR of the (i+1) x i Hessenberg matrix, built column by column
struct Givens {
    std::vector<double> c, s;
    void reset() {
        c.clear();
        s.clear();
    }
    // rotate a new column h[0..i] (in place) and append its rotation
    void push(std::vector<double>& h, int i) {
        for (int k = 0; k < i; k++) {
            double t = c[k] * h[k] + s[k] * h[k + 1];
            h[k + 1] = -s[k] * h[k] + c[k] * h[k + 1];
            h[k] = t;
        }
        double a = h[i], b = h[i + 1];
        double r = std::hypot(a, b);
        double cc = r == 0 ? 1.0 : a / r, ss = r == 0 ? 0.0 : b / r;
        c.push_back(cc);
        s.push_back(ss);
        h[i] = r;
        h[i + 1] = 0.0;
    }
};

// restarted GMRES(m) on A M^-1 y = b (M = ILU(0)), the caller's x = M^-1 y:
// at each restart r = b - A M^-1 y, eta = |r|/|r_first|; a full cycle of m
// Arnoldi steps (classical Gram-Schmidt) follows when eta > tol; the
// least-squares update is formed at the cycle's end.  Nothing inside a cycle
// depends on a host decision, so its Arnoldi steps run without a host round
// trip: each step's Gram-Schmidt coefficients and norm stay in a device
// column, and the host reads all m columns once, at the cycle's end, to apply
// the Givens rotations in step order.  Returns the number of Arnoldi steps; y
// is updated in place.
int IluSystem::gmres_dev(double* Y, const double* Bv, int m, double tol, int max_iter, double* eta_out) {
    if (!dperm_) return gmres_core(Y, Bv, m, tol, max_iter, eta_out);
    if (!ilu_ok_) throw NumericError("ILU(0): zero pivot");
    dev::permute(stream_, n_, dperm_, Y, vXo_, 0);
    dev::permute(stream_, n_, dperm_, Bv, vBo_, 0);
    const int it = gmres_core(vXo_, vBo_, m, tol, max_iter, eta_out);
    dev::permute(stream_, n_, dperm_, vXo_, Y, 1);
    hip_check(hipStreamSynchronize(stream_), "sync");
    return it;
}

int IluSystem::gmres_core(double* Y, const double* Bv, int m, double tol, int max_iter, double* eta_out) {
    if (!ilu_ok_) throw NumericError("ILU(0): zero pivot");
    // the multidot partials of a step (<= 512 blocks each) fit part_ up to m = 320
    if (m < 1 || m > 256) throw std::invalid_argument("GMRES restart length m must be in 1..256");
    hipStream_t st = stream_;
    const int n = n_;
    const int hl = m + 2;  // one column: coefficients 0..i-1, the norm at i
    if (mb_ < m + 2) {
        // the Krylov basis as one (m+2) x n block: V_l = V + l*n
        vbasis_ = kalloc<double>((size_t)(m + 2) * n, *bufs_);
        hcols_ = kalloc<double>((size_t)hl * hl, *bufs_);
        mb_ = m + 2;
    }
    double* V = vbasis_;
    const size_t ld = (size_t)n;
    double* coef = hcols_ + (size_t)m * hl;  // device copy of the update's coefficients (m <= 2 hl)
    double beta0 = 0, eta = 0;
    int iter = 0;
    std::vector<double> H((size_t)(m + 1) * m, 0.0), cols((size_t)hl * m), col(m + 2), g(m + 2), z(m + 1);
    Givens rot;
    for (;;) {
        // restart: r = b - A M^-1 y, eta = |r| / |r of the first restart|
        double* v0 = V;
        apply_p(Y, v0);
        matvec(v0, vW_);
        dev::axpby(st, n, 0.0, vR_, 1.0, Bv);
        dev::axpby(st, n, 1.0, vR_, -1.0, vW_);
        double beta = std::sqrt(dot(vR_, vR_));
        if (iter == 0) {
            beta0 = beta;
            eta = beta0 == 0.0 ? 0.0 : 1.0;
        } else {
            eta = beta / beta0;
        }
        if (eta > tol) {
            dev::axpby(st, n, 0.0, v0, 1.0 / beta, vR_);
            for (int i = 1; i <= m; i++) {
                double* vi = V + (size_t)(i - 1) * ld;
                double* vn = V + (size_t)i * ld;
                double* hc = hcols_ + (size_t)(i - 1) * hl;
                apply_p(vi, vn);
                matvec(vn, vR_);
                // classical Gram-Schmidt: all dots against the same r
                dev::multidot(st, n, i, V, ld, vR_, part_, hc);
                dev::combine(st, n, i, V, ld, hc, vR_, -1.0, vn);
                // |vn| into the column, then vn / |vn|
                const int np = dev::dot_part<double>(st, n, vn, vn, part_);
                dev::finalize(st, np, part_, hc, i, 1);
                dev::scale_by(st, n, vn, hc, i, 1);
                iter++;
            }
            hip_check(hipMemcpyAsync(cols.data(), hcols_, cols.size() * sizeof(double), hipMemcpyDeviceToHost, st),
                      "d2h");
            hip_check(hipStreamSynchronize(st), "sync");
            rot.reset();
            std::fill(g.begin(), g.end(), 0.0);
            g[0] = beta;
            for (int i = 1; i <= m; i++) {
                std::copy(cols.begin() + (size_t)(i - 1) * hl, cols.begin() + (size_t)(i - 1) * hl + i + 1,
                          col.begin());
                rot.push(col, i - 1);
                for (int k = 0; k <= i - 1; k++) H[(size_t)(i - 1) * (m + 1) + k] = col[k];
                double cc = rot.c[i - 1], ss = rot.s[i - 1];
                double t = cc * g[i - 1] + ss * g[i];
                g[i] = -ss * g[i - 1] + cc * g[i];
                g[i - 1] = t;
            }
            // least-squares update at the end of the cycle: R z = g, y += V z
            for (int r = m - 1; r >= 0; r--) {
                double s = g[r];
                for (int c = r + 1; c < m; c++) s -= H[(size_t)c * (m + 1) + r] * z[c];
                z[r] = s / H[(size_t)r * (m + 1) + r];
            }
            hip_check(hipMemcpyAsync(coef, z.data(), m * sizeof(double), hipMemcpyHostToDevice, st), "h2d");
            dev::combine(st, n, m, V, ld, coef, Y, 1.0, vY_);
            dev::copy(st, n, (const double*)vY_, Y);
        }
        if (!(eta > tol && iter < max_iter)) break;
    }
    hip_check(hipStreamSynchronize(st), "sync");
    if (eta_out) *eta_out = eta;
    return iter;
}

// FEM.c's solver(): H = x; gmres(H); x = precon(H)
int IluSystem::gmres_host(double* x, const double* b, int m, double tol, int max_iter, double* eta) {
    const size_t bytes = (size_t)n_ * sizeof(double);
    hip_check(hipMemcpyAsync(vX_, x, bytes, hipMemcpyHostToDevice, stream_), "h2d");
    hip_check(hipMemcpyAsync(vB_, b, bytes, hipMemcpyHostToDevice, stream_), "h2d");
    int it = gmres_dev(vX_, vB_, m, tol, max_iter, eta);
    apply(vX_, vZ_);
    hip_check(hipMemcpyAsync(x, vZ_, bytes, hipMemcpyDeviceToHost, stream_), "d2h");
    hip_check(hipStreamSynchronize(stream_), "sync");
    return it;
}

int IluSystem::pcg_host(double* x, double* b, double tol, int max_iter, int stop, double* res) {
    const size_t bytes = (size_t)n_ * sizeof(double);
    hip_check(hipMemcpyAsync(vX_, x, bytes, hipMemcpyHostToDevice, stream_), "h2d");
    hip_check(hipMemcpyAsync(vB_, b, bytes, hipMemcpyHostToDevice, stream_), "h2d");
    int it = pcg_dev(vX_, vB_, tol, max_iter, stop, res);
    hip_check(hipMemcpyAsync(x, vX_, bytes, hipMemcpyDeviceToHost, stream_), "d2h");
    hip_check(hipMemcpyAsync(b, vB_, bytes, hipMemcpyDeviceToHost, stream_), "d2h");
    hip_check(hipStreamSynchronize(stream_), "sync");
    return it;
}

void IluSystem::apply_host(const double* z, double* w) {
    const size_t bytes = (size_t)n_ * sizeof(double);
    hip_check(hipMemcpyAsync(vZ_, z, bytes, hipMemcpyHostToDevice, stream_), "h2d");
    apply(vZ_, vW_);
    hip_check(hipMemcpyAsync(w, vW_, bytes, hipMemcpyDeviceToHost, stream_), "d2h");
    hip_check(hipStreamSynchronize(stream_), "sync");
}

// -------------------------------------------------- BiCGStab (AMG precon)

double Solver::dot_host(const double* a, const double* b) {
    int np = dev::dot_part<double>(stream_, n_, a, b, part_);
    dev::finalize(stream_, np, part_, scal_, dev::S_T2, 0);
    return read_scalar(dev::S_T2);
}

// y = A x with the caller's (unflipped) finest matrix
void Solver::apply_A(const double* x, double* y) {
    DevLevel& top = L_.back();
    dev::spmv<double, double>(stream_, top.A, x, nullptr, y, dev::MODE_SET, nullptr, nullptr);
    if (has_neg_signs_) dev::signs(stream_, n_, signs_, y, y);
}

// BiCGStab_solve (BiCGStab.c:115-198); preconditioner = AMG_solve with this
// hierarchy, warm-started from the previous PP / PS as in the reference
int Solver::bicgstab_dev(double* X, const double* b, double eps, int max_iter, double* res) {
    hipStream_t st = stream_;
    const int n = n_;
    if (!bS_)
        for (double** p : {&bS_, &bPS_, &bPP_, &bAPP_, &bAPS_, &bR_, &bR0_, &bP_}) *p = dalloc_public(n);
    for (double* p : {bS_, bPS_, bPP_, bAPP_, bAPS_}) dev::fill<double>(st, n, p, 0.0);
    vcycles_ = 0;
    precon_calls_ = 0;
    // R = -(A X - b)
    apply_A(X, bR_);
    dev::axpby(st, n, 1.0, bR_, -1.0, b);
    dev::axpby(st, n, -1.0, bR_, 0.0, bR_);
    dev::copy(st, n, (const double*)bR_, bR0_);
    dev::copy(st, n, (const double*)bR_, bP_);
    int iter = 0;
    double r0 = std::sqrt(dot_host(bR_, bR_)), r = r0;
    if (r0 != 0) {
        do {
            amg_solve_dev(bP_, bPP_);
            apply_A(bPP_, bAPP_);
            double h = dot_host(bR_, bR0_);
            double t = dot_host(bAPP_, bR0_);
            if (t == 0) break;
            double alpha = h / t;
            dev::copy(st, n, (const double*)bR_, bS_);
            dev::axpby(st, n, 1.0, bS_, -alpha, bAPP_);
            amg_solve_dev(bS_, bPS_);
            apply_A(bPS_, bAPS_);
            double s = dot_host(bAPS_, bAPS_);
            if (s == 0) break;
            double omega = dot_host(bAPS_, bS_) / s;
            dev::axpby(st, n, 1.0, X, alpha, bPP_);
            dev::axpby(st, n, 1.0, X, omega, bPS_);
            dev::copy(st, n, (const double*)bS_, bR_);
            dev::axpby(st, n, 1.0, bR_, -omega, bAPS_);
            double beta = dot_host(bR_, bR0_) * alpha / (h * omega);
            dev::axpby(st, n, 1.0, bP_, -omega, bAPP_);
            dev::axpby(st, n, beta, bP_, 0.0, bP_);
            dev::axpby(st, n, 1.0, bP_, 1.0, bR_);
            r = std::sqrt(dot_host(bR_, bR_));
            iter++;
        } while (r / r0 > eps && iter <= max_iter);
    }
    if (res) *res = r0 == 0 ? 0.0 : std::fabs(r / r0);
    return iter;
}

int Solver::bicgstab_host(double* x, const double* b, double eps, int max_iter, double* res) {
    const size_t bytes = (size_t)n_ * sizeof(double);
    hip_check(hipMemcpyAsync(kX_, x, bytes, hipMemcpyHostToDevice, stream_), "h2d");
    hip_check(hipMemcpyAsync(kB_, b, bytes, hipMemcpyHostToDevice, stream_), "h2d");
    int it = bicgstab_dev(kX_, kB_, eps, max_iter, res);
    hip_check(hipMemcpyAsync(x, kX_, bytes, hipMemcpyDeviceToHost, stream_), "d2h");
    hip_check(hipStreamSynchronize(stream_), "sync");
    return it;
}

}  // namespace lina
