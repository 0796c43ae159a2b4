// Device-resident AMG hierarchy and the Krylov drivers of liblina_hip.
//
// The host setup (amg_setup.cpp) produces the hierarchy; this file uploads it
// once (CSR values in FP64 and an FP32 copy, LDS row-block schedules,
// colour permutations, level schedules) and runs the reference's solve
// algorithms with every vector operation on the device:
//
//   pcg        PCGS.c:163-233
//   AMG_solve  Multigrid.c:1863-1931
//   V-cycle    Multigrid.c:1536-1580
//
// The host only reads back the scalars that drive control flow (|R| per
// Krylov iteration, |f-Ax| per V-cycle).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/lina_hip.h"
#include "kernels.hpp"
#include "lina_internal.hpp"
#include "solver.hpp"

namespace lina {

// ------------------------------------------------------------ errors

static thread_local std::string g_err;

void set_error(const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
}

const char* last_error() { return g_err.c_str(); }

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw HipError(std::string(what) + ": " + hipGetErrorString(e));
}

// ------------------------------------------------------------ memory

DevBuffers::~DevBuffers() {
    for (void* p : ptrs) (void)hipFree(p);
}

// row-block schedule over rows [begin, end): at most max_nnz products and
// max_rows rows per block; a row longer than max_nnz gets a block alone.
// (Blocks of up to 1024 rows for the prolongations, a few entries a row,
// measured no faster.)
// Appends the blocks' first rows to starts.
void build_blocks(const Csr& A, int begin, int end, int max_nnz, std::vector<int>& starts, int max_rows) {
    int r = begin;
    int nnz = 0, rows = 0;
    if (begin >= end) return;
    starts.push_back(begin);
    while (r < end) {
        int len = A.rp[r + 1] - A.rp[r];
        if (rows > 0 && (nnz + len > max_nnz || rows == max_rows)) {
            starts.push_back(r);
            nnz = 0;
            rows = 0;
        }
        nnz += len;
        rows++;
        r++;
    }
}

// (first row, first entry) of each block start (the list ends with A.n)
std::vector<int2> block_pairs(const Csr& A, const std::vector<int>& starts_with_end) {
    std::vector<int2> bk(starts_with_end.size());
    for (size_t b = 0; b < starts_with_end.size(); b++) {
        bk[b].x = starts_with_end[b];
        bk[b].y = A.rp[starts_with_end[b]];
    }
    return bk;
}

// every index a kernel will dereference must be inside its array
static void validate(const Csr& A, const std::vector<int>& starts, int max_nnz) {
    if ((int)A.rp.size() != A.n + 1 || A.rp[0] != 0) throw std::logic_error("upload: bad row pointer");
    for (int i = 0; i < A.n; i++)
        if (A.rp[i + 1] < A.rp[i]) throw std::logic_error("upload: row pointer not monotone");
    for (int c : A.ci)
        if (c < 0 || c >= A.ncols) throw std::logic_error("upload: column index outside the matrix");
    for (size_t b = 0; b + 1 < starts.size(); b++) {
        int r0 = starts[b], r1 = starts[b + 1];
        if (r0 >= r1 || r1 > A.n) throw std::logic_error("upload: bad row block");
        if (r1 - r0 > dev::BLOCK_ROWS) throw std::logic_error("upload: row block too tall");
        if (A.rp[r1] - A.rp[r0] > max_nnz && r1 - r0 != 1) throw std::logic_error("upload: row block too wide");
    }
}

// Prolongation rows may carry sparse_zero's (i,0.0) seed at a column i that
// is not a coarse index (rows left empty by the reference); a zero product
// adds nothing, so such entries are dropped before upload.
static Csr drop_outside_zeros(const Csr& A) {
    Csr B;
    B.n = A.n;
    B.ncols = A.ncols;
    B.rp.assign(A.n + 1, 0);
    for (int i = 0; i < A.n; i++) {
        for (int p = A.rp[i]; p < A.rp[i + 1]; p++) {
            if (A.ci[p] >= A.ncols) {
                if (A.v[p] != 0.0) throw std::logic_error("upload: nonzero entry outside the matrix");
                continue;
            }
            B.ci.push_back(A.ci[p]);
            B.v.push_back(A.v[p]);
        }
        B.rp[i + 1] = (int)B.ci.size();
    }
    return B;
}

// 16-bit column codes of a schedule (see dev::DMat::cs): per block, the
// sorted columns are split at their largest gap into at most two clusters;
// when each spans < 32768 the block gets bases (b0, b1) and codes
// sel << 15 | (column - base_sel), otherwise b0 = -1 (32-bit columns).
// Returns the number of entries coded in 16 bits.
static long column_codes(const Csr& A, const std::vector<int>& starts_with_end, std::vector<uint16_t>& cs,
                         std::vector<int2>& cb) {
    const int nblk = (int)starts_with_end.size() - 1;
    cb.assign(std::max(nblk, 1), make_int2(-1, -1));
    cs.assign(A.ci.size(), 0);
    long coded = 0;
    std::vector<int> c;
    for (int b = 0; b < nblk; b++) {
        const int k0 = A.rp[starts_with_end[b]], k1 = A.rp[starts_with_end[b + 1]];
        if (k1 <= k0) continue;
        c.assign(A.ci.begin() + k0, A.ci.begin() + k1);
        std::sort(c.begin(), c.end());
        size_t g = 0;  // clusters [c0, c[g]] and [c[g+1], c.back()]
        for (size_t i = 1; i < c.size(); i++)
            if (c[i] - c[i - 1] > c[g + 1] - c[g]) g = i - 1;
        int b0 = c.front(), b1 = c.front();
        if (c.back() - c.front() > 32767) {
            if (c[g] - c.front() > 32767 || c.back() - c[g + 1] > 32767) continue;
            b1 = c[g + 1];
        }
        cb[b] = make_int2(b0, b1);
        for (int k = k0; k < k1; k++) {
            const int col = A.ci[k];
            cs[k] = (b1 != b0 && col >= b1) ? (uint16_t)(0x8000 | (col - b1)) : (uint16_t)(col - b0);
        }
        coded += k1 - k0;
    }
    return coded;
}

// share: a copy of A with other values uploaded before; its row pointers and
// column indices are reused, and its 16-bit offsets too when its schedule
// (share_starts) is the same
// first and last system (seg: a batch's row offsets per system) of each row
// block's rows; row r of A is row row_ids[r] of the level (ascending in a block)
static std::vector<int2> block_systems(const std::vector<int>& starts_with_end, const std::vector<int>& seg,
                                       const int* row_ids) {
    const int nblk = (int)starts_with_end.size() - 1;
    std::vector<int2> bs(std::max(nblk, 1), make_int2(0, 0));
    auto sys = [&](int r) {
        const int i = row_ids ? row_ids[r] : r;
        return (int)(std::upper_bound(seg.begin(), seg.end(), i) - seg.begin()) - 1;
    };
    for (int b = 0; b < nblk; b++) {
        const int r0 = starts_with_end[b], r1 = starts_with_end[b + 1] - 1;
        const int s0 = sys(r0), s1 = sys(r1);
        bs[b] = make_int2(std::min(s0, s1), std::max(s0, s1));
    }
    return bs;
}

template <typename VT, typename OT = double>
static void upload_mat(const Csr& A, const std::vector<int>& starts_with_end, dev::DMat<VT>& D, DevBuffers& B,
                       hipStream_t st, const dev::DMat<OT>* share = nullptr,
                       const std::vector<int>* share_starts = nullptr, const int* row_ids = nullptr,
                       const std::vector<int>* seg = nullptr) {
    validate(A, starts_with_end, dev::block_nnz<VT>());
    D.bspec = seg ? dupload(block_systems(starts_with_end, *seg, row_ids), B, st) : nullptr;
    D.active = nullptr;
    D.n = A.n;
    D.ncols = A.ncols;
    D.nnz = A.nnz();
    D.rp = share ? share->rp : dupload(A.rp, B, st);
    D.ci = share ? share->ci : dupload(A.ci, B, st);
    std::vector<VT> v(A.v.begin(), A.v.end());
    D.v = dupload(v, B, st);
    D.nblk = (int)starts_with_end.size() - 1;
    if (D.nblk < 0) D.nblk = 0;
    D.bk = dupload(block_pairs(A, starts_with_end), B, st);
    if (share && share_starts && *share_starts == starts_with_end) {
        D.cs = share->cs;
        D.cb = share->cb;
        D.idx_bytes = share->idx_bytes;
    } else {
        std::vector<uint16_t> cs;
        std::vector<int2> cb;
        const long coded = A.nnz() > 0 ? column_codes(A, starts_with_end, cs, cb) : 0;
        D.cs = nullptr;
        D.cb = nullptr;
        D.idx_bytes = 4.0 * A.nnz();
        if (coded > 0) {
            D.cs = dupload(cs, B, st);
            D.cb = dupload(cb, B, st);
            D.idx_bytes = 2.0 * coded + 4.0 * (A.nnz() - coded) + 8.0 * D.nblk;
        }
    }
    D.lpr = dev::lanes_per_row(A.n > 0 ? (double)A.nnz() / A.n : 0.0);
    // diagonal positions (row r's own index: row_ids == nullptr) and l1 norms
    std::vector<int> dpos(A.n, -1);
    std::vector<VT> l1(A.n, (VT)0);
    for (int r = 0; r < A.n; r++) {
        int i = row_ids ? row_ids[r] : r;
        VT s = 0;
        for (int p = A.rp[r]; p < A.rp[r + 1]; p++) {
            if (A.ci[p] == i) dpos[r] = p;
            s += std::fabs((VT)A.v[p]);
        }
        l1[r] = s;
    }
    D.dpos = dupload(dpos, B, st);
    D.l1 = dupload(l1, B, st);
}

// entries per row block of a matrix with nnz entries: the kernels' maximum,
// halved (down to 256) until the matrix has about 2048 blocks -- 8 per CU, so
// that the small coarse levels and transfers are not left to a few CUs
static int block_cap(long nnz, int cap) {
    int m = cap;
    while (m > 256 && nnz / m < 2048) m /= 2;
    return m;
}

// block schedule of [0, A.n) for value type VT, ending with A.n
template <typename VT>
static std::vector<int> schedule(const Csr& A) {
    std::vector<int> starts;
    build_blocks(A, 0, A.n, block_cap(A.nnz(), dev::block_nnz<VT>()), starts);
    starts.push_back(A.n);
    if (A.n == 0) starts.assign(1, 0);
    return starts;
}

// upload A as FP64 + FP32 copies sharing the index arrays (each with the row
// blocks of its own value size); returns the larger block count
int upload_pair(const Csr& A0, dev::DMat<double>& D, dev::DMat<float>& F, DevBuffers& B, hipStream_t st,
                const std::vector<int>* seg) {
    Csr A = drop_outside_zeros(A0);
    const std::vector<int> s64 = schedule<double>(A), s32 = schedule<float>(A);
    upload_mat<double, double>(A, s64, D, B, st, nullptr, nullptr, nullptr, seg);
    upload_mat(A, s32, F, B, st, &D, &s64, nullptr, seg);
    return std::max(D.nblk, F.nblk);
}

// level schedule of the lexicographic sweep: forward, level(i) = 1 + max over
// j<i with a_ij != 0; backward, over j>i
static void level_schedule(const Csr& A, bool backward, std::vector<int>& off, std::vector<int>& rows) {
    const int n = A.n;
    std::vector<int> lev(n, 0);
    int nl = 0;
    for (int t = 0; t < n; t++) {
        int i = backward ? n - 1 - t : t;
        int l = 0;
        for (int p = A.rp[i]; p < A.rp[i + 1]; p++) {
            int j = A.ci[p];
            if (A.v[p] == 0.0) continue;
            if ((!backward && j < i) || (backward && j > i)) l = std::max(l, lev[j] + 1);
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

// ------------------------------------------------------------ Solver

Solver::Solver() {}

Solver::~Solver() {
    if (stream_) (void)hipStreamSynchronize(stream_);
    if (ev0_) (void)hipEventDestroy(ev0_);
    if (ev1_) (void)hipEventDestroy(ev1_);
    for (auto& pr : evk_)
        for (hipEvent_t e : pr)
            if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : evr_)
        if (e) (void)hipEventDestroy(e);
    if (host_scal_) (void)hipHostFree(host_scal_);
    if (host_ring_) (void)hipHostFree(host_ring_);
    bufs_.reset();
    if (stream_) (void)hipStreamDestroy(stream_);
}

void Solver::setup_host(const Csr& A, const lina_amg_params& p) {
    prm_ = p;
    SetupParams sp;
    sp.dof = p.dof;
    sp.depth = p.depth;
    sp.smoothing_iterations = p.smoothing_iterations;
    sp.aggr_smoothing_iterations = p.aggr_smoothing_iterations;
    sp.coarsening = p.coarsening;
    sp.eps_minus = p.eps_minus;
    sp.eps_plus = p.eps_plus;
    sp.matrix_tol = p.matrix_tol;
    amg_setup(A, sp, H_);
    n_ = A.n;
    colours_.assign(H_.lv.size(), std::vector<int>());
    ncolours_.assign(H_.lv.size(), 0);
    for (size_t l = 1; l < H_.lv.size(); l++) colours_[l] = greedy_colouring(H_.lv[l].A, &ncolours_[l]);
}

// ------------------------------------------------------- autotuning

namespace {
constexpr int kLprCandidates[] = {1, 4, 8, 16, 32};
}

template <typename VT, typename XT>
int Solver::tune_spmv(dev::DMat<VT>& M, const XT* x, XT* y) {
    if (M.nblk == 0) return M.lpr;
    float best = 1e30f;
    int best_lpr = M.lpr;
    for (int lpr : kLprCandidates) {
        M.lpr = lpr;
        dev::spmv<VT, XT>(stream_, M, x, nullptr, y, dev::MODE_SET, nullptr, nullptr);
        hip_check(hipEventRecord(ev0_, stream_), "event");
        for (int k = 0; k < 3; k++) dev::spmv<VT, XT>(stream_, M, x, nullptr, y, dev::MODE_SET, nullptr, nullptr);
        hip_check(hipEventRecord(ev1_, stream_), "event");
        hip_check(hipEventSynchronize(ev1_), "event sync");
        float ms = 0;
        hip_check(hipEventElapsedTime(&ms, ev0_, ev1_), "elapsed");
        if (ms < best) {
            best = ms;
            best_lpr = lpr;
        }
    }
    M.lpr = best_lpr;
    return best_lpr;
}

template <typename VT, typename XT>
int Solver::tune_mcsor(DevLevel& d, dev::DMat<VT>& Ac, const XT* f, XT* x) {
    if (Ac.nblk == 0 || d.ncolours == 0) return Ac.lpr;
    float best = 1e30f;
    int best_lpr = Ac.lpr;
    const std::vector<int>& cb = sizeof(VT) == 8 ? d.colour_blk : d.colour_blk32;
    auto sweep = [&]() {
        for (int c = 0; c < d.ncolours; c++)
            dev::mcsor_colour(stream_, Ac, cb[c], cb[c + 1] - cb[c], d.perm, f, x, (XT)1);
    };
    for (int lpr : kLprCandidates) {
        Ac.lpr = lpr;
        sweep();
        hip_check(hipEventRecord(ev0_, stream_), "event");
        for (int k = 0; k < 2; k++) sweep();
        hip_check(hipEventRecord(ev1_, stream_), "event");
        hip_check(hipEventSynchronize(ev1_), "event sync");
        float ms = 0;
        hip_check(hipEventElapsedTime(&ms, ev0_, ev1_), "elapsed");
        if (ms < best) {
            best = ms;
            best_lpr = lpr;
        }
    }
    Ac.lpr = best_lpr;
    return best_lpr;
}

// The row partition (lanes per row in the LDS reduction phase) of every
// matrix of the hierarchy, chosen by timing each candidate on the device:
// the SpMV-shaped kernels (residual, restriction, prolongation, Jacobi) take
// their matrix's choice, the multicolour sweeps the colour-permuted copy's.
void Solver::autotune() {
    hipStream_t st = stream_;
    tuned_.clear();
    for (size_t l = 0; l < L_.size(); l++) {
        DevLevel& d = L_[l];
        for (double* v : {d.x, d.f, d.r, d.xo}) dev::fill<double>(st, d.n, v, 1.0);
        for (float* v : {d.xf, d.ff, d.rf, d.xof}) dev::fill<float>(st, d.n, v, 1.0f);
        tuned_.push_back({(int)l, 'A', tune_spmv<double, double>(d.A, d.x, d.r), tune_spmv<float, float>(d.Af, d.xf, d.rf)});
        if (l == 0) continue;
        DevLevel& c = L_[l - 1];
        for (double* v : {c.x, c.f}) dev::fill<double>(st, c.n, v, 1.0);
        for (float* v : {c.xf, c.ff}) dev::fill<float>(st, c.n, v, 1.0f);
        tuned_.push_back({(int)l, 'P', tune_spmv<double, double>(d.P, c.x, d.r), tune_spmv<float, float>(d.Pf, c.xf, d.rf)});
        tuned_.push_back({(int)l, 'R', tune_spmv<double, double>(d.R, d.r, c.f), tune_spmv<float, float>(d.Rf, d.rf, c.ff)});
        dev::fill<double>(st, d.n, d.f, 0.0);
        dev::fill<float>(st, d.n, d.ff, 0.0f);
        tuned_.push_back({(int)l, 'C', tune_mcsor<double, double>(d, d.Ac, d.f, d.x),
                          tune_mcsor<float, float>(d, d.Acf, d.ff, d.xf)});
    }
    hip_check(hipStreamSynchronize(st), "autotune");
}

void Solver::set_hierarchy(HostHierarchy&& H, const lina_amg_params& p) {
    prm_ = p;
    H_ = std::move(H);
    n_ = H_.lv.back().A.n;
    colours_.assign(H_.lv.size(), std::vector<int>());
    ncolours_.assign(H_.lv.size(), 0);
    for (size_t l = 1; l < H_.lv.size(); l++) colours_[l] = greedy_colouring(H_.lv[l].A, &ncolours_[l]);
}

void Solver::upload() {
    if (!stream_) {
        hip_check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate");
        hip_check(hipEventCreate(&ev0_), "hipEventCreate");
        hip_check(hipEventCreate(&ev1_), "hipEventCreate");
        hip_check(hipHostMalloc((void**)&host_scal_, dev::S_COUNT * sizeof(double)), "hipHostMalloc");
        hip_check(hipHostMalloc((void**)&host_ring_, 2 * dev::S_COUNT * sizeof(double)), "hipHostMalloc");
        for (hipEvent_t& e : evr_) hip_check(hipEventCreate(&e), "hipEventCreate");
    }
    // a re-upload (recompute) frees every earlier buffer, including the
    // BiCGStab work vectors allocated on first use
    hip_check(hipStreamSynchronize(stream_), "sync");
    if (bufs_ && (host_stale_ || host_top_stale_)) sync_host_levels();  // values of a device recompute
    rap_.reset();  // the device products' plans belong to the patterns uploaded before
    host_stale_ = false;
    host_top_stale_ = false;
    neg_rows_ = nullptr;
    iso_pos_ = nullptr;
    iso_val_ = nullptr;
    niso_pos_ = 0;
    iso_pos_host_.clear();
    bufs_.reset(new DevBuffers());
    for (double** p : {&bS_, &bPS_, &bPP_, &bAPP_, &bAPS_, &bR_, &bR0_, &bP_}) *p = nullptr;
    DevBuffers& B = *bufs_;
    hipStream_t st = stream_;
    const int depth = (int)H_.lv.size();
    L_.assign(depth, DevLevel());
    up_.assign(track_ ? depth : 0, UpLevel());
    int maxn = 1, maxblk = 1;
    for (int l = 0; l < depth; l++) {
        const HostLevel& h = H_.lv[l];
        DevLevel& d = L_[l];
        d.n = h.A.n;
        maxn = std::max(maxn, d.n);
        const std::vector<int>* segl = seg_off_.empty() ? nullptr : &seg_off_[l];
        maxblk = std::max(maxblk, upload_pair(h.A, d.A, d.Af, B, st, segl));
        d.x = dalloc<double>(d.n, B);
        d.f = dalloc<double>(d.n, B);
        d.r = dalloc<double>(d.n, B);
        d.xo = dalloc<double>(d.n, B);
        d.xf = dalloc<float>(d.n, B);
        d.ff = dalloc<float>(d.n, B);
        d.rf = dalloc<float>(d.n, B);
        d.xof = dalloc<float>(d.n, B);
        d.nlim = (d.n / std::max(1, prm_.dof)) * std::max(1, prm_.dof);
        if (track_) {
            up_[l].rp = h.A.rp;
            up_[l].ci = h.A.ci;
        }
        if (l == 0) continue;
        maxblk = std::max(maxblk, upload_pair(h.P, d.P, d.Pf, B, st, segl));
        maxblk = std::max(maxblk, upload_pair(h.R, d.R, d.Rf, B, st, seg_off_.empty() ? nullptr : &seg_off_[l - 1]));
        // colour-permuted copy for the multicolour sweeps
        const std::vector<int>& col = colours_[l];
        int nc = ncolours_[l];
        std::vector<int> perm(d.n);
        {
            std::vector<int> cnt(nc + 1, 0);
            for (int i = 0; i < d.n; i++) cnt[col[i] + 1]++;
            for (int c = 0; c < nc; c++) cnt[c + 1] += cnt[c];
            std::vector<int> fill(cnt.begin(), cnt.end() - 1);
            for (int i = 0; i < d.n; i++) perm[fill[col[i]]++] = i;
            d.colour_rows = cnt;
        }
        Csr Ac;
        Ac.n = d.n;
        Ac.ncols = d.n;
        Ac.rp.assign(d.n + 1, 0);
        Ac.ci.reserve(h.A.nnz());
        Ac.v.reserve(h.A.nnz());
        for (int r = 0; r < d.n; r++) {
            int i = perm[r];
            for (int p = h.A.rp[i]; p < h.A.rp[i + 1]; p++) {
                Ac.ci.push_back(h.A.ci[p]);
                Ac.v.push_back(h.A.v[p]);
            }
            Ac.rp[r + 1] = (int)Ac.ci.size();
        }
        // colour c's rows form the blocks [colour_blk[c], colour_blk[c+1])
        auto colour_schedule = [&](int max_nnz, std::vector<int>& cb) {
            std::vector<int> starts;
            cb.assign(nc + 1, 0);
            for (int c = 0; c < nc; c++) {
                cb[c] = (int)starts.size();
                build_blocks(Ac, d.colour_rows[c], d.colour_rows[c + 1], max_nnz, starts);
            }
            cb[nc] = (int)starts.size();
            starts.push_back(d.n);
            return starts;
        };
        const std::vector<int> c64 = colour_schedule(dev::block_nnz<double>(), d.colour_blk);
        const std::vector<int> c32 = colour_schedule(dev::block_nnz<float>(), d.colour_blk32);
        upload_mat<double, double>(Ac, c64, d.Ac, B, st, nullptr, nullptr, perm.data(), segl);
        upload_mat(Ac, c32, d.Acf, B, st, &d.Ac, &c64, perm.data(), segl);
        d.perm = dupload(perm, B, st);
        d.ncolours = nc;
        if (track_) {
            up_[l].colour = col;
            up_[l].perm = perm;
        }
        // level schedules of the exact lexicographic sweeps
        std::vector<int> off, rows;
        level_schedule(h.A, false, off, rows);
        d.nlv_f = (int)off.size() - 1;
        d.lvf_off = dupload(off, B, st);
        d.lvf_rows = dupload(rows, B, st);
        if (track_) {
            up_[l].lvf_off = off;
            up_[l].lvf_rows = rows;
        }
        level_schedule(h.A, true, off, rows);
        d.nlv_b = (int)off.size() - 1;
        d.lvb_off = dupload(off, B, st);
        d.lvb_rows = dupload(rows, B, st);
        if (track_) {
            up_[l].lvb_off = off;
            up_[l].lvb_rows = rows;
        }
        if (lex_smoothing()) build_sor_packs(l, B, false);
    }
    // finest-level extras
    signs_ = dupload(H_.signs, B, st);
    has_neg_signs_ = false;
    for (double s : H_.signs)
        if (s < 0) has_neg_signs_ = true;
    iso_rows_ = dupload(H_.iso_rows, B, st);
    iso_a_ = dupload(H_.iso_a, B, st);
    niso_ = (int)H_.iso_rows.size();
    // the isolated rows' diagonals by row for the fused preconditioner entry
    // and exit (a zero diagonal keeps the separate kernels)
    isa_ = nullptr;
    isb_ = nullptr;
    isa_ok_ = true;
    if (niso_ > 0) {
        std::vector<double> isa(n_, 0.0);
        std::vector<uint32_t> isb((n_ + 31) / 32, 0u);
        for (int k = 0; k < niso_; k++) {
            const int i = H_.iso_rows[k];
            if (H_.iso_a[k] == 0.0) isa_ok_ = false;
            isa[i] = H_.iso_a[k];
            isb[i >> 5] |= 1u << (i & 31);
        }
        if (isa_ok_) {
            isa_ = dupload(isa, B, st);
            isb_ = dupload(isb, B, st);
        }
    }
    // coarse LU inverse
    const CoarseLU& lu = H_.lu;
    nd_ = (int)lu.dense_idx.size();
    nt_ = (int)lu.trivial_idx.size();
    didx_ = dupload(lu.dense_idx, B, st);
    inv_ = dupload(lu.dense, B, st);
    cbatched_ = lu.blk_off.size() > 2;
    if (cbatched_) {
        // block-diagonal coarsest level of a batch: one inverse per block
        const int nb = (int)lu.blk_off.size() - 1;
        std::vector<int> rblk(nd_);
        int maxb = 0;
        for (int b = 0; b < nb; b++) {
            maxb = std::max(maxb, lu.blk_off[b + 1] - lu.blk_off[b]);
            for (int g = lu.blk_off[b]; g < lu.blk_off[b + 1]; g++) rblk[g] = b;
        }
        crblk_ = dupload(rblk, B, st);
        coff_ = dupload(lu.blk_off, B, st);
        cmoff_ = dupload(lu.mat_off, B, st);
        double* piv = dalloc<double>(nb, B);
        dev::batched_gauss_jordan_inverse(st, nd_, maxb, crblk_, coff_, cmoff_, inv_, piv);
    } else {
        cpiv_ = dalloc<double>(dev::gauss_jordan_work(nd_), B);
        dev::gauss_jordan_inverse(st, nd_, inv_, cpiv_);
    }
    up_dense_ = lu.dense_idx;
    up_trivial_ = lu.trivial_idx;
    // gathered right-hand side, padded with zeros to whole tiles of the FP32 inverse
    const int nbd = std::max((nd_ + dev::CT - 1) / dev::CT * dev::CT, dev::CT);
    cbd_ = dalloc<double>(nbd, B);
    dev::fill<double>(st, nbd, cbd_, 0.0);
    // dense index of each coarsest row (restriction epilogue) and whether the
    // dense and trivial rows cover the level (no zero fill needed)
    dinv_ = nullptr;
    ccover_ = false;
    if (!cbatched_) {
        const int n0 = L_[0].n;
        std::vector<int> dinv(n0, -1), seen(n0, 0);
        bool ok = true;
        for (int g = 0; g < nd_ && ok; g++) {
            const int r = lu.dense_idx[g];
            if (r < 0 || r >= n0 || seen[r]++) ok = false;
            else dinv[r] = g;
        }
        for (int t = 0; t < nt_ && ok; t++) {
            const int r = lu.trivial_idx[t];
            if (r < 0 || r >= n0 || seen[r]++) ok = false;
        }
        if (ok) {
            dinv_ = dupload(dinv, B, st);
            ccover_ = nd_ + nt_ == n0;
        }
    }
    ctiles_ = nullptr;
    nct_ = ntiles_ = 0;
    tidx_ = dupload(lu.trivial_idx, B, st);
    tdiag_ = dupload(lu.trivial_diag, B, st);
    // Krylov vectors and scratch
    const int n = n_;
    ftop_ = dalloc<double>(n, B);
    rtop_ = dalloc<double>(n, B);
    kR_ = dalloc<double>(n, B);
    kP_ = dalloc<double>(n, B);
    kQ_ = dalloc<double>(n, B);
    kZ_ = dalloc<double>(n, B);
    kW_ = dalloc<double>(n, B);
    kX_ = dalloc<double>(n, B);
    kB_ = dalloc<double>(n, B);
    // block partials: one per SpMV block (any matrix) or per reduction block
    int np = std::max(4096, maxblk + 16);
    part_ = dalloc<double>(np, B);
    part_cap_ = np;
    scal_ = dalloc<double>(dev::S_COUNT, B);
    dev::fill(st, dev::S_COUNT, scal_, 0.0);
    hip_check(hipStreamSynchronize(st), "upload sync");
    ensure_ctiles();
    if (prm_.autotune) autotune();
}

void Solver::set_params(const lina_amg_params& p) { prm_ = p; }

void Solver::set_segments(std::vector<std::vector<int>> off) {
    if (!off.empty() && off.size() != H_.lv.size()) throw std::invalid_argument("segments: one offset list per level");
    seg_off_ = std::move(off);
}

void Solver::set_active(const double* active) {
    if (seg_off_.empty()) throw std::logic_error("set_active without segments");
    for (DevLevel& d : L_) {
        for (auto* m : {&d.A, &d.P, &d.R, &d.Ac}) m->active = m->bspec ? active : nullptr;
        for (auto* m : {&d.Af, &d.Pf, &d.Rf, &d.Acf}) m->active = m->bspec ? active : nullptr;
    }
    seg_active_ = active;
}

// The coarse inverse's upper triangle as FP64 tiles of CT x CT, read by
// both precisions' V-cycles (half the bytes of the full inverse, exactly
// symmetric).  FP32 tiles were measured: 35 % more CG iterations on the
// high-contrast adjoint system, none on the elasticity benchmark.
void Solver::ensure_ctiles() {
    if (ctiles_ || cbatched_ || nd_ == 0 || !uploaded()) return;
    nct_ = (nd_ + dev::CT - 1) / dev::CT;
    std::vector<int> ti, tj;
    for (int I = 0; I < nct_; I++)
        for (int J = I; J < nct_; J++) {
            ti.push_back(I);
            tj.push_back(J);
        }
    ntiles_ = (int)ti.size();
    cti_ = dupload(ti, *bufs_, stream_);
    ctj_ = dupload(tj, *bufs_, stream_);
    cpart_ = dalloc<double>((size_t)nct_ * nct_ * dev::CT, *bufs_);
    ctiles_ = dalloc<double>((size_t)ntiles_ * dev::CT * dev::CT, *bufs_);
    dev::inverse_to_tiles<double>(stream_, nd_, ntiles_, cti_, ctj_, inv_, ctiles_);
    hip_check(hipStreamSynchronize(stream_), "sync");
}

// whether any level sweeps with the exact lexicographic SOR / SSOR
bool Solver::lex_smoothing() const {
    auto lex = [](int sm) { return sm == LINA_SMOOTHER_SOR || sm == LINA_SMOOTHER_SSOR; };
    return lex(prm_.smoother) || lex(prm_.coarse_smoother);
}

// The level-packed forms (dev::SorPack) of level l's exact lexicographic
// sweeps, from the host matrix and the level schedules: new entries (earlier
// rows in the sweep's order, nonzero, as the schedule counts them) in the
// records (past 4 or 8 in overflow arrays), old entries (all, zero or not)
// in the slot-major lists and their overflow.  Not built (the level keeps
// k_sor_lex) when a row lacks its diagonal.  reuse: the device arrays of the
// level's packs are rewritten in place when their sizes still fit.
void Solver::build_sor_packs(int l, DevBuffers& B, bool reuse) {
    DevLevel& d = L_[l];
    reuse = reuse && d.sor_packed;
    d.sor_packed = false;
    const Csr& A = H_.lv[l].A;
    const int n = A.n;
    if (n == 0) return;
    hipStream_t st = stream_;
    std::vector<int> off, rows;
    int maxc = 0;
    for (int i = 0; i < n; i++) {
        int nl = 0, nu = 0, dg = 0;
        for (int p = A.rp[i]; p < A.rp[i + 1]; p++) {
            const int j = A.ci[p];
            if (j == i) dg++;
            else if (j < i) nl++;
            else nu++;
        }
        if (dg != 1) return;
        maxc = std::max(maxc, std::max(nl, nu));
    }
    const int maxp = maxc <= 4 ? 4 : 8;
    for (int back = 0; back < 2; back++) {
        level_schedule(A, back != 0, off, rows);
        const int nlv = (int)off.size() - 1;
        std::vector<int> pos(n), lev(n);
        for (int L = 0; L < nlv; L++)
            for (int t = off[L]; t < off[L + 1]; t++) {
                pos[rows[t]] = t;
                lev[rows[t]] = L;
            }
        const size_t N = (size_t)n;
        std::vector<int> full(std::max(nlv, 1), 0), ocnt(n), ocol(maxp * N, 0), ovp(n, 0), ovsrc, ooff(n, 0), oovc;
        std::vector<double> oval(maxp * N, 0.0), ovval, oovv;
        std::vector<char> rec(N * (maxp == 4 ? sizeof(dev::TriRec<4>) : sizeof(dev::TriRec<8>)));
        for (int t = 0; t < n; t++) {
            const int i = rows[t], L = lev[i], e = off[L + 1];
            double diag = 0;
            double val[8];
            int src[8];
            for (int k = 0; k < 8; k++) {
                val[k] = 0.0;
                src[k] = dev::TRI_RING;
            }
            int kn = 0, ko = 0;
            ovp[t] = (int)ovval.size();
            ooff[t] = (int)oovv.size();
            for (int p = A.rp[i]; p < A.rp[i + 1]; p++) {
                const int j = A.ci[p];
                if (j == i) {
                    diag = A.v[p];
                } else if ((j < i) != (back != 0)) {  // new value
                    if (A.v[p] == 0.0) continue;
                    int sj;
                    if (e - pos[j] <= dev::TRI_RING) {
                        sj = pos[j] & (dev::TRI_RING - 1);
                    } else {
                        sj = -(j + 1);
                        full[L] = 1;
                    }
                    if (kn < maxp) {
                        val[kn] = A.v[p];
                        src[kn] = sj;
                    } else {
                        ovval.push_back(A.v[p]);
                        ovsrc.push_back(sj);
                    }
                    kn++;
                } else {  // old value
                    if (ko < maxp) {
                        oval[ko * N + t] = A.v[p];
                        ocol[ko * N + t] = j;
                    } else {
                        oovv.push_back(A.v[p]);
                        oovc.push_back(j);
                    }
                    ko++;
                }
            }
            ocnt[t] = ko;
            if (maxp == 4) {
                dev::TriRec<4>& r = reinterpret_cast<dev::TriRec<4>*>(rec.data())[t];
                r.diag = diag;
                for (int k = 0; k < 4; k++) {
                    r.val[k] = val[k];
                    r.src[k] = src[k];
                }
                r.orow = i;
                r.cnt = kn;
            } else {
                dev::TriRec<8>& r = reinterpret_cast<dev::TriRec<8>*>(rec.data())[t];
                r.diag = diag;
                for (int k = 0; k < 8; k++) {
                    r.val[k] = val[k];
                    r.src[k] = src[k];
                }
                r.orow = i;
                r.cnt = kn;
            }
        }
        for (int L = 0; L < nlv; L++)
            if (full[L]) off[L] |= (int)0x80000000u;
        dev::SorPack& S = back ? d.sorb : d.sorf;
        std::vector<int>& used = back ? d.sor_sizes_b : d.sor_sizes_f;
        const std::vector<int> sizes = {(int)ovval.size(), (int)oovv.size()};
        if (reuse && used == sizes) {
            // same pattern and schedules (checked by the caller) and overflow sizes
            auto to = [&](const void* dst, const void* src, size_t bytes) {
                if (bytes)
                    hip_check(hipMemcpyAsync(const_cast<void*>(dst), src, bytes, hipMemcpyHostToDevice, st), "h2d");
            };
            to(S.p.off, off.data(), off.size() * sizeof(int));
            to(S.p.rec, rec.data(), rec.size());
            to(S.p.ovp, ovp.data(), ovp.size() * sizeof(int));
            to(S.p.ovval, ovval.data(), ovval.size() * sizeof(double));
            to(S.p.ovsrc, ovsrc.data(), ovsrc.size() * sizeof(int));
            to(S.row, rows.data(), rows.size() * sizeof(int));
            to(S.ocnt, ocnt.data(), ocnt.size() * sizeof(int));
            to(S.oval, oval.data(), oval.size() * sizeof(double));
            to(S.ocol, ocol.data(), ocol.size() * sizeof(int));
            to(S.ooff, ooff.data(), ooff.size() * sizeof(int));
            to(S.oovv, oovv.data(), oovv.size() * sizeof(double));
            to(S.oovc, oovc.data(), oovc.size() * sizeof(int));
            hip_check(hipStreamSynchronize(st), "sync");  // the host arrays die here
            continue;
        }
        used = sizes;
        S = dev::SorPack();
        S.p.n = n;
        S.p.nlv = nlv;
        S.p.maxp = maxp;
        S.p.off = dupload(off, B, st);
        S.p.rec = dupload(rec, B, st);
        S.p.ovp = dupload(ovp, B, st);
        S.p.ovval = dupload(ovval, B, st);
        S.p.ovsrc = dupload(ovsrc, B, st);
        S.row = dupload(rows, B, st);
        S.ocnt = dupload(ocnt, B, st);
        S.oval = dupload(oval, B, st);
        S.ocol = dupload(ocol, B, st);
        S.ooff = dupload(ooff, B, st);
        S.oovv = dupload(oovv, B, st);
        S.oovc = dupload(oovc, B, st);
        S.old_first = back;
        S.in = dalloc<double>(n, B);
        S.a = dalloc<double>(n, B);
        S.prod = dalloc<double>(maxp * N, B);
        S.prodov = dalloc<double>(std::max<size_t>(oovv.size(), 1), B);
        dev::fill<double>(st, (int)(maxp * N), S.prod, 0.0);
        hip_check(hipStreamSynchronize(st), "sync");  // the host arrays die here
    }
    d.sor_packed = true;
}

// host l1 norms of a matrix's rows as upload_mat forms them (VT arithmetic)
template <typename VT>
static std::vector<VT> row_l1(const Csr& A) {
    std::vector<VT> l1(A.n, (VT)0);
    for (int r = 0; r < A.n; r++) {
        VT s = 0;
        for (int p = A.rp[r]; p < A.rp[r + 1]; p++) s += std::fabs((VT)A.v[p]);
        l1[r] = s;
    }
    return l1;
}

void Solver::refresh_after_recompute() {
    const bool timing = getenv("LINA_RN_TIMING") != nullptr;  // read per call: recomputes are rare
    const auto t0 = std::chrono::steady_clock::now();
    for (size_t l = 1; l < H_.lv.size(); l++) colours_[l] = greedy_colouring(H_.lv[l].A, &ncolours_[l]);
    const auto t1 = std::chrono::steady_clock::now();
    bool vals = false;
    if (uploaded()) {
        vals = refresh_values() && refresh_isolated();
        if (!vals) {
            track_ = true;  // from now on upload() keeps what a values-only refresh compares
            upload();
        }
    } else {
        refresh_isolated();
    }
    if (timing) {
        const auto t2 = std::chrono::steady_clock::now();
        auto s = [](auto a, auto b) { return std::chrono::duration<double>(b - a).count(); };
        fprintf(stderr, "[amg] recompute: colouring %.3fs, %s upload %.3fs\n", s(t0, t1), vals ? "values" : "full",
                s(t1, t2));
    }
}

// The values-only re-upload after a recompute: every level's A (FP64, FP32,
// the colour-permuted copies and their row l1 norms) and the coarse inverse
// and tiles, when each level's pattern, colouring and lexicographic level
// schedules and the coarse LU's row split are those on the device (P, R and
// the autotuned lane counts stay).  False, with nothing changed, otherwise.
bool Solver::refresh_values() {
    const int depth = (int)H_.lv.size();
    if ((int)up_.size() != depth || (int)L_.size() != depth || cbatched_ || !seg_off_.empty()) return false;
    for (int l = 0; l < depth; l++) {
        const Csr& A = H_.lv[l].A;
        const UpLevel& u = up_[l];
        if (A.rp != u.rp || A.ci != u.ci || A.nnz() != L_[l].A.nnz) return false;
        if (l == 0) continue;
        if (colours_[l] != u.colour) return false;
        std::vector<int> off, rows;
        level_schedule(A, false, off, rows);
        if (off != u.lvf_off || rows != u.lvf_rows) return false;
        level_schedule(A, true, off, rows);
        if (off != u.lvb_off || rows != u.lvb_rows) return false;
    }
    const CoarseLU& lu = H_.lu;
    if (lu.dense_idx != up_dense_ || lu.trivial_idx != up_trivial_ || !cpiv_) return false;
    hipStream_t st = stream_;
    std::vector<std::vector<double>> keep;  // host staging alive until the copies are done
    std::vector<std::vector<float>> keepf;
    auto put = [&](double* d, std::vector<double> h) {
        keep.push_back(std::move(h));
        hip_check(hipMemcpyAsync(d, keep.back().data(), keep.back().size() * sizeof(double), hipMemcpyHostToDevice,
                                 st), "h2d");
    };
    auto putf = [&](float* d, std::vector<float> h) {
        keepf.push_back(std::move(h));
        hip_check(hipMemcpyAsync(d, keepf.back().data(), keepf.back().size() * sizeof(float), hipMemcpyHostToDevice,
                                 st), "h2d");
    };
    for (int l = 0; l < depth; l++) {
        const Csr& A = H_.lv[l].A;
        DevLevel& d = L_[l];
        put(d.A.v, A.v);
        dev::convert<double, float>(st, d.A.nnz, d.A.v, d.Af.v);
        put(d.A.l1, row_l1<double>(A));
        putf(d.Af.l1, row_l1<float>(A));
        if (l == 0) continue;
        const std::vector<int>& perm = up_[l].perm;
        Csr Ac;
        Ac.n = A.n;
        Ac.rp.assign(A.n + 1, 0);
        Ac.v.reserve(A.nnz());
        for (int r = 0; r < A.n; r++) {
            const int i = perm[r];
            Ac.v.insert(Ac.v.end(), A.v.begin() + A.rp[i], A.v.begin() + A.rp[i + 1]);
            Ac.rp[r + 1] = (int)Ac.v.size();
        }
        put(d.Ac.l1, row_l1<double>(Ac));
        putf(d.Acf.l1, row_l1<float>(Ac));
        const int nnz = (int)Ac.v.size();
        put(d.Ac.v, std::move(Ac.v));
        dev::convert<double, float>(st, nnz, d.Ac.v, d.Acf.v);
        if (d.sor_packed) {
            hip_check(hipStreamSynchronize(st), "sync");
            build_sor_packs(l, *bufs_, true);  // the packs hold values: rewritten in place
        }
    }
    put(inv_, lu.dense);
    dev::gauss_jordan_inverse(st, nd_, inv_, cpiv_);
    put(tdiag_, lu.trivial_diag);
    if (ctiles_) dev::inverse_to_tiles<double>(st, nd_, ntiles_, cti_, ctj_, inv_, ctiles_);
    hip_check(hipStreamSynchronize(st), "refresh sync");
    return true;
}

void Solver::recompute(const std::vector<double>& val) {
    Csr A = H_.lv.back().A;  // pattern of the finest (sign-adjusted) matrix
    if ((int)val.size() != A.nnz()) throw std::invalid_argument("recompute: value count differs from the pattern");
    for (int i = 0; i < A.n; i++)
        for (int p = A.rp[i]; p < A.rp[i + 1]; p++) A.v[p] = val[p] * (H_.signs[i] < 0 ? -1.0 : 1.0);
    const bool timing = getenv("LINA_RN_TIMING") != nullptr;  // read per call: recomputes are rare
    if (device_rap_ok()) {
        const auto t0 = std::chrono::steady_clock::now();
        const bool built = rap_ != nullptr;
        const bool ok = recompute_device(A);
        if (timing) {
            const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            fprintf(stderr, "[amg] recompute on the device: %s %.3fs%s\n", ok ? "done" : "pattern changed", s,
                    built ? "" : " (plan built)");
        }
        if (ok) {
            H_.lv.back().A = std::move(A);
            host_stale_ = true;
            if (!refresh_isolated()) {
                sync_host_levels();
                upload();
            }
            return;
        }
    }
    sync_host_levels();
    amg_recompute(H_, A);
    refresh_after_recompute();
}

void Solver::recompute_matrix(const Csr& M) {
    if (M.n != n_) throw std::invalid_argument("recompute: matrix size differs from the hierarchy");
    const Csr& F = H_.lv.back().A;
    if (M.rp == F.rp && M.ci == F.ci) {  // the finest pattern holds: the values-only path
        recompute(M.v);
        return;
    }
    Csr A = M;
    for (int i = 0; i < A.n; i++)
        if (H_.signs[i] < 0)
            for (int p = A.rp[i]; p < A.rp[i + 1]; p++) A.v[p] = -A.v[p];
    sync_host_levels();
    amg_recompute(H_, A);
    refresh_after_recompute();
}

// AMG_fix_isolated (Multigrid.c:1301-1319) divides by the first nonzero of
// each isolated finest row: new finest values change the divisors
bool Solver::refresh_isolated() {
    if (host_top_stale_ && !H_.iso_rows.empty()) fetch_isolated_rows();
    const Csr& A = H_.lv.back().A;
    bool zero = false;
    for (size_t k = 0; k < H_.iso_rows.size(); k++) {
        const int i = H_.iso_rows[k];
        double a = 0.0;
        for (int p = A.rp[i]; p < A.rp[i + 1]; p++)
            if (A.v[p] != 0.0) {
                a = A.v[p];
                break;
            }
        H_.iso_a[k] = a;
        zero = zero || a == 0.0;
    }
    if (!uploaded() || niso_ == 0) return true;
    if (zero && isa_ok_) return false;  // the fused entry/exit kernels need nonzero divisors
    hip_check(hipMemcpyAsync(iso_a_, H_.iso_a.data(), (size_t)niso_ * sizeof(double), hipMemcpyHostToDevice, stream_),
              "h2d");
    if (isa_) {
        std::vector<double> isa(n_, 0.0);
        for (int k = 0; k < niso_; k++) isa[H_.iso_rows[k]] = H_.iso_a[k];
        hip_check(hipMemcpyAsync(isa_, isa.data(), (size_t)n_ * sizeof(double), hipMemcpyHostToDevice, stream_),
                  "h2d");
    }
    hip_check(hipStreamSynchronize(stream_), "sync");
    return true;
}

void Solver::set_operator(DevMatrix* op) {
    if (op) {
        if (op->n() != n_) throw std::invalid_argument("operator size differs from the hierarchy");
        if (op->nblocks() + 16 > part_cap_) throw std::invalid_argument("operator has more row blocks than the "
                                                                        "solver's partial-sum buffer");
        // the operator's own stream must be idle: its kernels now run on ours
        hip_check(hipStreamSynchronize(op->stream()), "sync");
    }
    op_ = op;
}

void Solver::set_kernel_timing(bool on) {
    if (on && !evk_[0][0])
        for (auto& pr : evk_)
            for (hipEvent_t& e : pr) hip_check(hipEventCreate(&e), "hipEventCreate");
    ktime_ = on;
}

double* Solver::dalloc_public(size_t n) { return dalloc<double>(n, *bufs_); }

// dL/dtheta of P1 plane-strain elements (host arrays in and out)
void Solver::elasticity_gradient(int n_nodes, const double* xy, int n_elem, const int* tri, double nu,
                                 const double* scale, int blocked, const double* u, const double* lam,
                                 double* grad) {
    DevBuffers B;
    auto up = [&](const void* h, size_t bytes) {
        void* d = nullptr;
        hip_check(hipMalloc(&d, std::max<size_t>(bytes, 8)), "hipMalloc");
        B.ptrs.push_back(d);
        if (h && bytes) hip_check(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, stream_), "h2d");
        return d;
    };
    const int n = 2 * n_nodes;
    auto* dxy = (const double*)up(xy, 2 * (size_t)n_nodes * sizeof(double));
    auto* dtri = (const int*)up(tri, 3 * (size_t)n_elem * sizeof(int));
    auto* du = (const double*)up(u, (size_t)n * sizeof(double));
    auto* dl = (const double*)up(lam, (size_t)n * sizeof(double));
    auto* ds = scale ? (const double*)up(scale, (size_t)n_elem * sizeof(double)) : nullptr;
    auto* dg = (double*)up(nullptr, (size_t)n_elem * sizeof(double));
    dev::elast_grad(stream_, n_elem, dtri, dxy, n_nodes, du, dl, nu, ds, blocked, dg);
    hip_check(hipMemcpyAsync(grad, dg, (size_t)n_elem * sizeof(double), hipMemcpyDeviceToHost, stream_), "d2h");
    hip_check(hipStreamSynchronize(stream_), "sync");
}

double Solver::read_scalar(int slot) {
    hip_check(hipMemcpyAsync(host_scal_ + slot, scal_ + slot, sizeof(double), hipMemcpyDeviceToHost, stream_), "d2h");
    hip_check(hipStreamSynchronize(stream_), "sync");
    return host_scal_[slot];
}

// ----------------------------------------------------------- smoothing

template <typename T>
static inline const dev::DMat<T>& pickA(const DevLevel& d);
template <>
inline const dev::DMat<double>& pickA<double>(const DevLevel& d) { return d.A; }
template <>
inline const dev::DMat<float>& pickA<float>(const DevLevel& d) { return d.Af; }
template <typename T>
static inline const dev::DMat<T>& pickAc(const DevLevel& d);
template <>
inline const dev::DMat<double>& pickAc<double>(const DevLevel& d) { return d.Ac; }
template <>
inline const dev::DMat<float>& pickAc<float>(const DevLevel& d) { return d.Acf; }
template <typename T>
static inline const dev::DMat<T>& pickP(const DevLevel& d);
template <>
inline const dev::DMat<double>& pickP<double>(const DevLevel& d) { return d.P; }
template <>
inline const dev::DMat<float>& pickP<float>(const DevLevel& d) { return d.Pf; }
template <typename T>
static inline const dev::DMat<T>& pickR(const DevLevel& d);
template <>
inline const dev::DMat<double>& pickR<double>(const DevLevel& d) { return d.R; }
template <>
inline const dev::DMat<float>& pickR<float>(const DevLevel& d) { return d.Rf; }

template <typename T>
struct LevelVecs;
template <>
struct LevelVecs<double> {
    static double* x(DevLevel& d) { return d.x; }
    static double* f(DevLevel& d) { return d.f; }
    static double* r(DevLevel& d) { return d.r; }
    static double* xo(DevLevel& d) { return d.xo; }
};
template <>
struct LevelVecs<float> {
    static float* x(DevLevel& d) { return d.xf; }
    static float* f(DevLevel& d) { return d.ff; }
    static float* r(DevLevel& d) { return d.rf; }
    static float* xo(DevLevel& d) { return d.xof; }
};

// smoother and sweep count in effect on level l (coarse_smoother overrides
// the levels below the finest)
void Solver::level_smoother(int l, int& smoother, int& sweeps) const {
    smoother = prm_.smoother;
    if (l < (int)L_.size() - 1 && prm_.coarse_smoother >= 0) {
        smoother = prm_.coarse_smoother;
        sweeps = prm_.coarse_sweeps;
    }
    if (l > 0 && l < (int)L_.size() - 1 && l < 32 && ((prm_.skip_levels >> l) & 1u)) sweeps = 0;
}

// smoother kind and sweep count of level l's pre- and post-smoothing in the
// V-cycle (AMG_V_cycle's Iter[level] settings)
void Solver::vcycle_smoother(int l, int& smoother, int& sweeps) const {
    const int depth = (int)L_.size();
    sweeps = (prm_.coarsening == AGGRESSIVE && l == depth - 1) ? prm_.aggr_smoothing_iterations
                                                                : prm_.smoothing_iterations;
    level_smoother(l, smoother, sweeps);
}

static bool is_jacobi(int smoother) {
    return smoother == LINA_SMOOTHER_JACOBI || smoother == LINA_SMOOTHER_L1_JACOBI;
}

// xinit (Solver::XInit): X_GUESS: x holds the iterate; X_ZERO: x stands for
// the zero vector and is not initialised (a coarse level's guess in the
// V-cycle), the smoother writes every entry; X_J0: x holds the first Jacobi
// sweep from zero (written by the restriction's epilogue); X_SPARE: the
// iterate is in the level's spare vector (a prolongation written beside x).
// X_J0 and X_SPARE are for Jacobi smoothers only.
template <typename T>
void Solver::smooth(int l, const T* f, T* x, int sweeps, bool post, int xinit) {
    DevLevel& d = L_[l];
    hipStream_t st = stream_;
    const T omega = (T)prm_.omega;
    int smoother;
    level_smoother(l, smoother, sweeps);
    const bool jac = is_jacobi(smoother);
    if ((xinit == X_J0 || xinit == X_SPARE) && !(jac && sweeps > 0))
        throw std::logic_error("smooth: a fused first step without a Jacobi sweep");
    const bool mc = smoother == LINA_SMOOTHER_MC_SOR || smoother == LINA_SMOOTHER_MC_SSOR;
    // a forward multicolour sweep from zero starts with one elementwise pass
    // (the first colour's values, zeros elsewhere); a batch's skipped row
    // blocks keep the zero fill instead
    const bool mc_first = mc && xinit == X_ZERO && sweeps > 0 && d.ncolours > 0 &&
                          !(smoother == LINA_SMOOTHER_MC_SSOR && post) && pickAc<T>(d).bspec == nullptr;
    if (xinit == X_ZERO && !(jac && sweeps > 0) && !mc_first) dev::fill<T>(st, d.n, x, (T)0);
    switch (smoother) {
        case LINA_SMOOTHER_SOR:
        case LINA_SMOOTHER_SSOR: {
            bool back = (smoother == LINA_SMOOTHER_SSOR) && post;
            if (std::is_same<T, double>::value && d.sor_packed)
                dev::sor_packed(st, back ? d.sorb : d.sorf, (const double*)f, (double*)x, (double)omega, sweeps);
            else if (back)
                dev::sor_lex(st, pickA<T>(d), d.nlv_b, d.lvb_off, d.lvb_rows, f, x, LevelVecs<T>::xo(d), omega,
                             sweeps, 1);
            else
                dev::sor_lex(st, pickA<T>(d), d.nlv_f, d.lvf_off, d.lvf_rows, f, x, LevelVecs<T>::xo(d), omega,
                             sweeps, 0);
            break;
        }
        case LINA_SMOOTHER_MC_SOR:
        case LINA_SMOOTHER_MC_SSOR: {
            bool back = (smoother == LINA_SMOOTHER_MC_SSOR) && post;
            const std::vector<int>& cb = sizeof(T) == 8 ? d.colour_blk : d.colour_blk32;
            for (int s = 0; s < sweeps; s++)
                for (int cc = 0; cc < d.ncolours; cc++) {
                    int c = back ? d.ncolours - 1 - cc : cc;
                    if (s == 0 && cc == 0 && mc_first) {
                        dev::mcsor_first(st, pickAc<T>(d), d.colour_rows[1], d.perm, f, x, omega);
                        continue;
                    }
                    int b0 = cb[c], b1 = cb[c + 1];
                    dev::mcsor_colour(st, pickAc<T>(d), b0, b1 - b0, d.perm, f, x, omega);
                }
            break;
        }
        case LINA_SMOOTHER_JACOBI:
        case LINA_SMOOTHER_L1_JACOBI: {
            // ping-pong between x and the level's spare vector, ending in x
            // when the iterate starts where the caller put it for that
            // (otherwise one copy back).  From a zero guess the first sweep
            // is elementwise.
            int kind = smoother == LINA_SMOOTHER_JACOBI ? dev::JAC_REF : dev::JAC_L1;
            T* a = x;
            T* b = LevelVecs<T>::xo(d);
            int s = 0;
            if (xinit == X_ZERO && sweeps > 0) {
                dev::jacobi0(st, pickA<T>(d), kind, f, x, d.nlim);
                s = 1;
            } else if (xinit == X_J0) {
                s = 1;
            } else if (xinit == X_SPARE) {
                std::swap(a, b);
            }
            for (; s < sweeps; s++) {
                dev::jacobi(st, pickA<T>(d), kind, f, (const T*)a, b, d.nlim);
                std::swap(a, b);
            }
            if (a != x) dev::copy(st, d.n, (const T*)a, x);
            break;
        }
        default:
            throw std::invalid_argument("unknown smoother");
    }
}

// the coarsest level's dense right-hand side is gathered by the restriction
bool Solver::fused_gather() const { return !cbatched_ && nd_ > 0 && dinv_ != nullptr && L_.size() > 1; }

// AMG_V_cycle (Multigrid.c:1536-1580) with Iter[level] == 1.  xinit as for
// smooth: X_ZERO stands for the reference zeroing the coarse guess before the
// recursive call.  Two steps are fused into neighbouring kernels where the
// arithmetic allows: a coarse level's first Jacobi sweep from zero (and the
// coarsest level's gather) into the restriction's epilogue, and a Jacobi
// post-smoother with an odd sweep count reads the prolongation from the
// spare vector, so its last sweep lands in x without a copy.
template <typename T>
void Solver::vcycle(int l, const T* f, T* x, int xinit) {
    hipStream_t st = stream_;
    if (l == 0) {
        // the dense-row/trivial-row split need not cover every row
        if (xinit == X_ZERO && !ccover_) dev::fill<T>(st, L_[0].n, x, (T)0);
        if (cbatched_)
            dev::bcoarse_solve(st, nd_, crblk_, coff_, cmoff_, didx_, inv_, nt_, tidx_, tdiag_, f, x, seg_active_);
        else if (ctiles_)
            dev::coarse_solve_sym(st, nd_, nct_, ntiles_, cti_, ctj_, (const double*)ctiles_, cpart_, didx_, nt_,
                                  tidx_, tdiag_, f, x, cbd_, fused_gather());
        else
            dev::coarse_solve(st, nd_, didx_, inv_, nt_, tidx_, tdiag_, f, x, cbd_, fused_gather());
        return;
    }
    DevLevel& d = L_[l];
    DevLevel& c = L_[l - 1];
    int sm, sw;
    vcycle_smoother(l, sm, sw);
    // a level without smoothing (skip_levels) passes its right-hand side down
    // and the coarse correction up: x = P V(R f), no matrix of its own read
    const bool pass = sw == 0 && xinit == X_ZERO && l < (int)L_.size() - 1;
    T* r = LevelVecs<T>::r(d);
    if (!pass) {
        smooth<T>(l, f, x, sw, false, xinit);
        dev::spmv<T, T>(st, pickA<T>(d), x, f, r, dev::MODE_RES, nullptr, nullptr);
    }
    T* fc = LevelVecs<T>::f(c);
    T* xc = LevelVecs<T>::x(c);
    dev::Epi<T, T> e;
    bool epi = false;
    int cinit = X_ZERO;
    if (l - 1 == 0) {
        if (fused_gather()) {
            e.bd = cbd_;
            e.dinv = dinv_;
            epi = true;
        }
    } else {
        int csm, csw;
        vcycle_smoother(l - 1, csm, csw);
        if (is_jacobi(csm) && csw > 0) {
            const auto& Ac = pickA<T>(c);
            e.x0 = xc;
            e.dpos = Ac.dpos;
            e.dv = Ac.v;
            e.l1 = Ac.l1;
            e.nlim = c.nlim;
            e.kind = csm == LINA_SMOOTHER_JACOBI ? dev::JAC_REF : dev::JAC_L1;
            epi = true;
            cinit = X_J0;
        }
    }
    dev::spmv<T, T>(st, pickR<T>(d), pass ? f : r, nullptr, fc, epi ? dev::MODE_SETX : dev::MODE_SET, nullptr,
                    nullptr, epi ? &e : nullptr);
    vcycle<T>(l - 1, fc, xc, cinit);
    if (pass) {
        dev::spmv<T, T>(st, pickP<T>(d), xc, nullptr, x, dev::MODE_SET, nullptr, nullptr);
        return;
    }
    if (is_jacobi(sm) && sw % 2 == 1) {
        dev::spmv<T, T>(st, pickP<T>(d), xc, x, LevelVecs<T>::xo(d), dev::MODE_ADDB, nullptr, nullptr);
        smooth<T>(l, f, x, sw, true, X_SPARE);
    } else {
        dev::spmv<T, T>(st, pickP<T>(d), xc, nullptr, x, dev::MODE_ADD, nullptr, nullptr);
        smooth<T>(l, f, x, sw, true, X_GUESS);
    }
}

// AMG_full_multigrid / FMG_recursive (Multigrid.c:1582-1611) with the
// recursion unrolled: the right-hand side restricted to every level
// (sparse_mult of each restriction, the reference's CF), the coarsest level
// solved, then on each finer level x = P x_coarse (sparse_mult, replacing
// the level's iterate) followed by one V-cycle.  The reference also
// restricts the incoming iterate down the levels (Coarse_X = R X), but the
// coarsest solve overwrites it and every finer level's iterate is replaced
// by the prolongation, so it never reaches the result and is not formed.
// Each level's f/x vectors hold CF[l] and the level's iterate: the V-cycle
// on level l writes only the vectors of the levels below, which the pass
// no longer reads.
template <typename T>
void Solver::fmg(const T* f, T* x) {
    hipStream_t st = stream_;
    const int top = (int)L_.size() - 1;
    for (int l = top; l >= 1; l--) {
        DevLevel& d = L_[l];
        const T* fl = l == top ? f : LevelVecs<T>::f(d);
        // the coarsest level's gathered right-hand side comes from the
        // restriction's epilogue when the coarse solve expects it there
        dev::Epi<T, T> e;
        const bool epi = l == 1 && fused_gather();
        if (epi) {
            e.bd = cbd_;
            e.dinv = dinv_;
        }
        dev::spmv<T, T>(st, pickR<T>(d), fl, nullptr, LevelVecs<T>::f(L_[l - 1]),
                        epi ? dev::MODE_SETX : dev::MODE_SET, nullptr, nullptr, epi ? &e : nullptr);
    }
    vcycle<T>(0, LevelVecs<T>::f(L_[0]), LevelVecs<T>::x(L_[0]), X_ZERO);
    for (int l = 1; l <= top; l++) {
        DevLevel& d = L_[l];
        T* xl = l == top ? x : LevelVecs<T>::x(d);
        const T* fl = l == top ? f : LevelVecs<T>::f(d);
        dev::spmv<T, T>(st, pickP<T>(d), LevelVecs<T>::x(L_[l - 1]), nullptr, xl, dev::MODE_SET, nullptr, nullptr);
        vcycle<T>(l, fl, xl, X_GUESS);
    }
}

// |f - A x| on the finest level into scal[slot] (FP64 or the FP32 copy)
template <typename T>
double Solver::residual_norm(const T* f, const T* x, T* r, int slot) {
    DevLevel& d = L_.back();
    const auto& A = pickA<T>(d);
    int np = dev::spmv<T, T>(stream_, A, x, f, r, dev::MODE_RES, part_, nullptr);
    dev::finalize(stream_, np, part_, scal_, slot, 1);
    return read_scalar(slot);
}

// AMG_solve (Multigrid.c:1863-1931) with F, X device vectors of the finest
// size.  The diagnostic ILU residuals of the reference (res0, res) do not
// influence the result except through `res0 != 0`, which holds exactly when
// the first residual r0 is nonzero; they are not computed.  LINA_CYCLE_FULL
// (the FULL_AMG case): one full-multigrid pass, counted as one cycle.
int Solver::amg_solve_dev(const double* F, double* X, double* relres) {
    hipStream_t st = stream_;
    DevLevel& top = L_.back();
    const int n = top.n;
    const double* Ft = F;
    if (has_neg_signs_) {
        dev::signs(st, n, signs_, F, ftop_);
        Ft = ftop_;
    }
    dev::fix_isolated<double>(st, niso_, iso_rows_, iso_a_, Ft, X);
    int it = 0;
    // relres: |r|/|r0| after the last V-cycle (AMG_add_stat's record,
    // Multigrid.c:1918); 0 when r0 = 0, 1 after a full-multigrid pass (the
    // reference leaves r = r0 there).  The loop skips the residual after the
    // last permitted cycle; it is formed only when asked for.
    double rel = 0.0;
    if (prm_.precision == LINA_PREC_FP64) {
        double r0 = residual_norm<double>(Ft, X, rtop_, dev::S_T0);
        if (r0 == 0.0) {
            if (relres) *relres = 0.0;
            return 0;
        }
        double r = r0;
        if (prm_.cycle == LINA_CYCLE_FULL) {
            fmg<double>(Ft, X);
            it = 1;
        }
        else while (std::fabs(r / r0) > prm_.amg_eps && it < prm_.amg_max_iter) {
            vcycle<double>((int)L_.size() - 1, Ft, X);
            it++;
            if (it < prm_.amg_max_iter || relres) r = residual_norm<double>(Ft, X, rtop_, dev::S_T0);
        }
        rel = std::fabs(r / r0);
    } else {
        float* ff = top.ff;
        float* xf = top.xf;
        dev::convert<double, float>(st, n, Ft, ff);
        dev::convert<double, float>(st, n, X, xf);
        double r0 = residual_norm<float>(ff, xf, top.rf, dev::S_T0);
        if (r0 != 0.0) {
            double r = r0;
            if (prm_.cycle == LINA_CYCLE_FULL) {
                fmg<float>(ff, xf);
                it = 1;
            }
            else while (std::fabs(r / r0) > prm_.amg_eps && it < prm_.amg_max_iter) {
                vcycle<float>((int)L_.size() - 1, ff, xf);
                it++;
                if (it < prm_.amg_max_iter || relres) r = residual_norm<float>(ff, xf, top.rf, dev::S_T0);
            }
            dev::convert<float, double>(st, n, xf, X);
            rel = std::fabs(r / r0);
        }
    }
    if (relres) *relres = rel;
    vcycles_ += it;
    precon_calls_++;
    return it;
}

// AMG_solve from a zero guess with at most one V-cycle, without a host round
// trip.  The reference's first residual is r0 = |F - A*0| = |F|; only
// `r0 == 0` matters with one V-cycle (the loop test is r/r0 = 1 > eps
// otherwise), and it holds exactly when every F_i^2 underflows, in whatever
// order the squares are summed.  The V-cycle runs unconditionally and its
// result is replaced by the untouched guess on the device when r0 == 0.
// w != nullptr: also the block partials of w.X into part_ (their number in
// *np), the dot k_dot_part(w, X) would give.
// prepared: pcg_xr already wrote this entry's outputs for F (the signed
// right-hand side, its FP32 copy, the initial guess, scal[S_AMG0]); see
// entry_prepared.
int Solver::vcycle_once(const double* F, double* X, const double* w, int* np, bool prepared) {
    hipStream_t st = stream_;
    DevLevel& top = L_.back();
    const int n = top.n;
    const int lt = (int)L_.size() - 1;
    const double* Ft = F;
    if (has_neg_signs_) {
        if (!prepared) dev::signs(st, n, signs_, F, ftop_);
        Ft = ftop_;
    }
    // without isolated rows the guess is exactly zero and the pre-smoother
    // writes it; otherwise AMG_fix_isolated sets those rows first
    const bool xz = niso_ == 0;
    const bool fp64 = prm_.precision == LINA_PREC_FP64;
    int npd = 0;
    if (isa_ok_) {
        // entry and exit fused into one elementwise kernel each (the entry
        // into pcg_xr when prepared)
        if (!prepared) {
            int n0;
            if (fp64) n0 = dev::precon_in<double>(st, n, Ft, part_, xz ? nullptr : X, nullptr, nullptr, isb_, isa_);
            else n0 = dev::precon_in<float>(st, n, Ft, part_, nullptr, xz ? nullptr : top.xf, top.ff, isb_, isa_);
            dev::finalize(st, n0, part_, scal_, dev::S_AMG0, 0);
        }
        if (fp64) {
            vcycle<double>(lt, Ft, X, xz ? X_ZERO : X_GUESS);
            npd = dev::precon_out<double>(st, n, nullptr, X, Ft, isb_, isa_, scal_, dev::S_AMG0, w, part_);
        } else {
            vcycle<float>(lt, top.ff, top.xf, xz ? X_ZERO : X_GUESS);
            npd = dev::precon_out<float>(st, n, top.xf, X, Ft, isb_, isa_, scal_, dev::S_AMG0, w, part_);
        }
    } else {
        if (prepared) throw std::logic_error("vcycle_once: prepared entry without the fused path");
        int n0 = dev::dot_part<double>(st, n, Ft, Ft, part_);
        dev::finalize(st, n0, part_, scal_, dev::S_AMG0, 0);
        if (!xz) {
            dev::fill<double>(st, n, X, 0.0);
            dev::fix_isolated<double>(st, niso_, iso_rows_, iso_a_, Ft, X);
        }
        if (fp64) {
            vcycle<double>(lt, Ft, X, xz ? X_ZERO : X_GUESS);
        } else {
            dev::convert<double, float>(st, n, Ft, top.ff);
            if (!xz) dev::convert<double, float>(st, n, X, top.xf);
            vcycle<float>(lt, top.ff, top.xf, xz ? X_ZERO : X_GUESS);
            dev::convert<float, double>(st, n, top.xf, X);
        }
        dev::restore_if_zero(st, n, X, scal_, dev::S_AMG0, niso_, iso_rows_, iso_a_, Ft);
        if (w) npd = dev::dot_part<double>(st, n, w, X, part_);
    }
    if (np) *np = npd;
    vcycles_ += 1;
    precon_calls_++;
    return 1;
}

// the preconditioner runs without host round trips (pcg's iterations can then
// be queued ahead of the stop test)
bool Solver::precon_async() const {
    return prm_.zero_init && prm_.amg_max_iter == 1 && prm_.cycle == LINA_CYCLE_V;
}

// pcg_xr can write the next preconditioner entry (vcycle_once's fused path;
// not with make_positive_diag signs, whose signed right-hand side precon_out
// would also need)
bool Solver::entry_prepared_ok() const { return precon_async() && isa_ok_ && !has_neg_signs_; }

int Solver::precon(const double* R, double* Z, const double* w, int* np, bool prepared) {
    if (precon_async()) return vcycle_once(R, Z, w, np, prepared);
    if (prepared) throw std::logic_error("precon: prepared entry without a one-V-cycle preconditioner");
    if (prm_.zero_init) dev::fill<double>(stream_, n_, Z, 0.0);
    int it = amg_solve_dev(R, Z);
    if (w) *np = dev::dot_part<double>(stream_, n_, w, Z, part_);
    return it;
}

// pcg (PCGS.c:163-233) on device vectors X, B (length n).  The loop test is
// evaluated on the device (k_finalize_stop); with a preconditioner that needs
// no host round trip, iteration k+1 is queued before the host waits for
// iteration k's test, and an iteration queued after the stop leaves X, R and
// the scalars untouched.
int Solver::pcg_dev(double* X, double* B, double tol, int max_iter, int stop, lina_stats* stats) {
    hipStream_t st = stream_;
    const int n = n_;
    DevLevel& top = L_.back();
    vcycles_ = 0;
    precon_calls_ = 0;
    hip_check(hipEventRecord(ev0_, st), "event");
    dev::fill<double>(st, 2, scal_ + dev::S_STOP, 0.0);  // S_STOP, S_IT
    // h = |B| (h==0 -> 1), B and X scaled by 1/h
    int np = dev::dot_part<double>(st, n, B, B, part_);
    dev::finalize(st, np, part_, scal_, dev::S_H, 1);
    dev::scale_by(st, n, B, scal_, dev::S_H, 1);
    dev::scale_by(st, n, X, scal_, dev::S_H, 1);
    double* R = kR_;
    double* P = kP_;
    double* Q = kQ_;
    double* Z = kZ_;
    double* W = kW_;
    dev::fill<double>(st, n, W, 0.0);
    dev::fill<double>(st, n, Z, 0.0);
    // R = B + (-1)*(A X) with the caller's matrix (rows of the stored finest
    // matrix carry the make_positive_diag signs)
    if (op_) {
        op_->launch_on(st, X, B, R, dev::MODE_RES, nullptr, nullptr);
    } else if (has_neg_signs_) {
        apply_A(X, Q);
        dev::copy(st, n, (const double*)B, R);
        dev::axpby(st, n, 1.0, R, -1.0, Q);
    } else {
        dev::spmv<double, double>(st, top.A, X, B, R, dev::MODE_RES, nullptr, nullptr);
    }
    precon(R, W, W, &np);
    dev::finalize(st, np, part_, scal_, dev::S_R0, 1);
    double r0 = read_scalar(dev::S_R0);
    if (r0 == 0) {
        // the reference returns here without undoing the 1/h scaling
        if (stats) {
            stats->iterations = 0;
            stats->residual = 0;
        }
        return 0;
    }
    const int relres = stop == LINA_STOP_RELRES;
    const int cap = std::max(max_iter, 1);  // the loop body runs at least once
    // one iteration: Z = M^-1 R with R.Z, P, Q = A P with P.Q, X and R with
    // |R| and the stop test; its state (S_R..S_IT) copied to ring slot k % 2
    // kernel timing samples every fourth iteration's outer SpMV (the events
    // themselves cost a few microseconds)
    auto sampled = [&](int k) { return ktime_ && k % 4 == 0; };
    // from the second iteration on, pcg_xr writes the preconditioner's entry
    // on the new R (its FP32 copy, initial guess and r0 test)
    const bool prep = entry_prepared_ok();
    auto enqueue = [&](int k) {
        int m = 0;
        precon(R, Z, R, &m, prep && k > 0);
        dev::finalize(st, m, part_, scal_, dev::S_S, 2);  // s_old = s; s = R.Z
        dev::pcg_p(st, n, P, Z, scal_, k == 0);
        if (op_) {
            m = op_->launch_on(st, P, nullptr, Q, dev::MODE_OUTER, part_, P);
        } else if (has_neg_signs_) {
            apply_A(P, Q);
            m = dev::dot_part<double>(st, n, P, Q, part_);
        } else {
            if (sampled(k))  // the events bracket the kernel itself
                m = dev::spmv_outer_timed(st, top.A, P, Q, part_, P, evk_[k % 2][0], evk_[k % 2][1]);
            else
                m = dev::spmv<double, double>(st, top.A, P, nullptr, Q, dev::MODE_OUTER, part_, P);  // Q = A P, P.Q
        }
        dev::finalize(st, m, part_, scal_, dev::S_PQ, 0);
        if (!prep) {
            m = dev::pcg_xr(st, n, X, R, P, Q, scal_, part_);
        } else if (prm_.precision == LINA_PREC_FP64) {
            m = dev::pcg_xr<double>(st, n, X, R, P, Q, scal_, part_, nullptr, niso_ == 0 ? nullptr : Z, nullptr,
                                    isb_, isa_);
        } else {
            m = dev::pcg_xr<float>(st, n, X, R, P, Q, scal_, part_, top.ff, niso_ == 0 ? nullptr : top.xf, nullptr,
                                   isb_, isa_);
        }
        dev::finalize_stop(st, m, part_, scal_, tol, relres, cap);
        double* h = host_ring_ + (k % 2) * dev::S_COUNT;
        hip_check(hipMemcpyAsync(h + dev::S_R, scal_ + dev::S_R, (dev::S_IT - dev::S_R + 1) * sizeof(double),
                                 hipMemcpyDeviceToHost, st),
                  "d2h");
        hip_check(hipEventRecord(evr_[k % 2], st), "event");
    };
    const bool timed = ktime_ && !has_neg_signs_ && !op_;
    const int ahead = precon_async() ? 2 : 1;
    double ksum = 0.0;
    int kcount = 0;
    int enq = 0, done = 0;
    double r = 0.0;
    int i = 0;
    for (;;) {
        while (enq < cap && enq - done < ahead) enqueue(enq++);
        hip_check(hipEventSynchronize(evr_[done % 2]), "event sync");
        const double* h = host_ring_ + (done % 2) * dev::S_COUNT;
        const int it_dev = (int)h[dev::S_IT];
        if (timed && sampled(done) && it_dev == done + 1) {
            float kms = 0;
            hip_check(hipEventElapsedTime(&kms, evk_[done % 2][0], evk_[done % 2][1]), "elapsed");
            ksum += kms;
            kcount++;
        }
        r = h[dev::S_R];
        i = it_dev;
        done++;
        if (h[dev::S_STOP] != 0.0 || (done == enq && enq >= cap)) break;
    }
    // an iteration queued after the stop ran its V-cycle and nothing else
    vcycles_ -= enq - i;
    precon_calls_ -= enq - i;
    double relres_out = r;
    precon(R, W, W, &np);
    dev::finalize(st, np, part_, scal_, dev::S_T1, 1);
    double rw = read_scalar(dev::S_T1);
    dev::scale_by(st, n, B, scal_, dev::S_H, 0);
    dev::scale_by(st, n, X, scal_, dev::S_H, 0);
    hip_check(hipEventRecord(ev1_, st), "event");
    hip_check(hipEventSynchronize(ev1_), "event sync");
    float ms = 0;
    hip_check(hipEventElapsedTime(&ms, ev0_, ev1_), "elapsed");
    pcg_residuum_ = rw / r0;
    if (stats) {
        stats->iterations = i;
        stats->precon_calls = precon_calls_;
        stats->vcycles = vcycles_;
        stats->residual = pcg_residuum_;
        stats->relres = relres_out;
        stats->t_solve_ms = ms;
        stats->spmv_ms = kcount ? ksum / kcount : 0.0;
        stats->spmv_launches = kcount;
    }
    if (std::isnan(rw / r0)) throw NumericError("pcg: NaN residual");
    return i;
}

int Solver::pcg_host(double* x, double* b, double tol, int max_iter, int stop, lina_stats* stats) {
    const size_t bytes = (size_t)n_ * sizeof(double);
    hip_check(hipMemcpyAsync(kX_, x, bytes, hipMemcpyHostToDevice, stream_), "h2d");
    hip_check(hipMemcpyAsync(kB_, b, bytes, hipMemcpyHostToDevice, stream_), "h2d");
    int it = pcg_dev(kX_, kB_, tol, max_iter, stop, stats);
    hip_check(hipMemcpyAsync(x, kX_, bytes, hipMemcpyDeviceToHost, stream_), "d2h");
    hip_check(hipMemcpyAsync(b, kB_, bytes, hipMemcpyDeviceToHost, stream_), "d2h");
    hip_check(hipStreamSynchronize(stream_), "sync");
    return it;
}

int Solver::amg_solve_host(const double* f, double* x, double* relres) {
    const size_t bytes = (size_t)n_ * sizeof(double);
    hip_check(hipMemcpyAsync(kB_, f, bytes, hipMemcpyHostToDevice, stream_), "h2d");
    hip_check(hipMemcpyAsync(kX_, x, bytes, hipMemcpyHostToDevice, stream_), "h2d");
    vcycles_ = 0;
    precon_calls_ = 0;
    int it = amg_solve_dev(kB_, kX_, relres);
    hip_check(hipMemcpyAsync(x, kX_, bytes, hipMemcpyDeviceToHost, stream_), "d2h");
    hip_check(hipStreamSynchronize(stream_), "sync");
    return it;
}

int Solver::spmv_host(int level, int precision, const double* x, double* y) {
    DevLevel& d = L_.at(level);
    const size_t bytes = (size_t)d.n * sizeof(double);
    hip_check(hipMemcpyAsync(d.xo, x, bytes, hipMemcpyHostToDevice, stream_), "h2d");
    if (precision == LINA_PREC_FP64) {
        dev::spmv<double, double>(stream_, d.A, d.xo, nullptr, d.r, dev::MODE_SET, nullptr, nullptr);
    } else {
        dev::convert<double, float>(stream_, d.n, d.xo, d.xof);
        dev::spmv<float, float>(stream_, d.Af, d.xof, nullptr, d.rf, dev::MODE_SET, nullptr, nullptr);
        dev::convert<float, double>(stream_, d.n, d.rf, d.r);
    }
    hip_check(hipMemcpyAsync(y, d.r, bytes, hipMemcpyDeviceToHost, stream_), "d2h");
    hip_check(hipStreamSynchronize(stream_), "sync");
    return 0;
}

double Solver::time_spmv(int level, int precision, int reps) {
    DevLevel& d = L_.at(level);
    dev::fill<double>(stream_, d.n, d.xo, 1.0);
    dev::fill<float>(stream_, d.n, d.xof, 1.0f);
    auto launch = [&]() {
        if (precision == LINA_PREC_FP64)
            dev::spmv<double, double>(stream_, d.A, d.xo, nullptr, d.r, dev::MODE_OUTER, nullptr, nullptr);
        else
            dev::spmv<float, float>(stream_, d.Af, d.xof, nullptr, d.rf, dev::MODE_OUTER, nullptr, nullptr);
    };
    for (int i = 0; i < 3; i++) launch();
    hip_check(hipEventRecord(ev0_, stream_), "event");
    for (int i = 0; i < reps; i++) launch();
    hip_check(hipEventRecord(ev1_, stream_), "event");
    hip_check(hipEventSynchronize(ev1_), "event sync");
    float ms = 0;
    hip_check(hipEventElapsedTime(&ms, ev0_, ev1_), "elapsed");
    return ms / reps;
}

double spmv_bytes(double n, double nnz, int precision) {
    if (precision == LINA_PREC_FP64) return 12.0 * nnz + 4.0 * (n + 1) + 16.0 * n;
    return 8.0 * nnz + 4.0 * (n + 1) + 8.0 * n;
}

// algorithmic bytes of one y = A x with the uploaded arrays: values, column
// indices (2-byte codes + 8 bytes of bases per block, or 4 bytes), row
// pointers, block pairs, x gathered once, y written once
template <typename VT>
static double mat_bytes(const dev::DMat<VT>& A) {
    return (double)sizeof(VT) * A.nnz + A.idx_bytes + 4.0 * (A.n + 1) + 8.0 * (A.nblk + 1) +
           (double)sizeof(VT) * (A.ncols + A.n);
}

double Solver::spmv_bytes(int level, int precision) const {
    if (uploaded() && level >= 0 && level < (int)L_.size())
        return precision == LINA_PREC_FP64 ? mat_bytes(L_[level].A) : mat_bytes(L_[level].Af);
    const Csr& A = H_.lv.at(level).A;
    return lina::spmv_bytes(A.n, A.nnz(), precision);
}

// ------------------------------------------------------- DevMatrix

DevMatrix::DevMatrix(const Csr& A) : n_(A.n), nnz_(A.nnz()), bufs_(new DevBuffers) {
    hip_check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "stream");
    hip_check(hipEventCreate(&ev0_), "event");
    hip_check(hipEventCreate(&ev1_), "event");
    upload_pair(A, A_, Af_, *bufs_, stream_);
    prp_ = A.rp;
    pci_ = A.ci;
    x_ = dalloc<double>(n_, *bufs_);
    y_ = dalloc<double>(n_, *bufs_);
    xf_ = dalloc<float>(n_, *bufs_);
    yf_ = dalloc<float>(n_, *bufs_);
    hip_check(hipStreamSynchronize(stream_), "upload");
}

bool DevMatrix::update_values(const Csr& A0) {
    if (A0.n != n_ || A0.rp != prp_ || A0.ci != pci_) return false;
    const Csr A = drop_outside_zeros(A0);
    if (A.nnz() != A_.nnz) return false;
    hip_check(hipMemcpyAsync(A_.v, A.v.data(), A.v.size() * sizeof(double), hipMemcpyHostToDevice, stream_), "h2d");
    dev::convert<double, float>(stream_, A_.nnz, A_.v, Af_.v);
    hip_check(hipStreamSynchronize(stream_), "update");
    return true;
}

bool DevMatrix::update_values_device(const double* d_val, long nnz) {
    if (nnz != (long)A_.nnz || nnz_ != nnz) return false;
    hip_check(hipMemcpyAsync(A_.v, d_val, nnz * sizeof(double), hipMemcpyDeviceToDevice, stream_), "d2d");
    dev::convert<double, float>(stream_, A_.nnz, A_.v, Af_.v);
    hip_check(hipStreamSynchronize(stream_), "update");
    return true;
}

double DevMatrix::matrix_norm() const {
    std::vector<double> v(A_.nnz);
    hip_check(hipMemcpyAsync(v.data(), A_.v, v.size() * sizeof(double), hipMemcpyDeviceToHost, stream_), "d2h");
    hip_check(hipStreamSynchronize(stream_), "sync");
    double s = 0;
    for (double a : v) s += std::fabs(a);
    return s / n_;
}

DevMatrix::~DevMatrix() {
    if (stream_) (void)hipStreamSynchronize(stream_);
    bufs_.reset();
    if (ev0_) (void)hipEventDestroy(ev0_);
    if (ev1_) (void)hipEventDestroy(ev1_);
    if (stream_) (void)hipStreamDestroy(stream_);
}

void DevMatrix::apply(const double* x, double* y) {
    dev::spmv<double, double>(stream_, A_, x, nullptr, y, dev::MODE_OUTER, nullptr, nullptr);
}

int DevMatrix::launch(const double* x, const double* b, double* y, int mode, double* part, const double* w) {
    return dev::spmv<double, double>(stream_, A_, x, b, y, mode, part, w);
}

int DevMatrix::launch_on(hipStream_t st, const double* x, const double* b, double* y, int mode, double* part,
                         const double* w) {
    return dev::spmv<double, double>(st, A_, x, b, y, mode, part, w);
}

void DevMatrix::spmv_dev(const double* x, double* y) {
    apply(x, y);
    hip_check(hipStreamSynchronize(stream_), "spmv");
}

void DevMatrix::spmv_host(int precision, const double* x, double* y) {
    const size_t bytes = (size_t)n_ * sizeof(double);
    hip_check(hipMemcpyAsync(x_, x, bytes, hipMemcpyHostToDevice, stream_), "h2d");
    if (precision == LINA_PREC_FP64) {
        apply(x_, y_);
    } else {
        dev::convert<double, float>(stream_, n_, x_, xf_);
        dev::spmv<float, float>(stream_, Af_, xf_, nullptr, yf_, dev::MODE_OUTER, nullptr, nullptr);
        dev::convert<float, double>(stream_, n_, yf_, y_);
    }
    hip_check(hipMemcpyAsync(y, y_, bytes, hipMemcpyDeviceToHost, stream_), "d2h");
    hip_check(hipStreamSynchronize(stream_), "sync");
}

double DevMatrix::time_spmv(int precision, int reps) {
    dev::fill<double>(stream_, n_, x_, 1.0);
    dev::fill<float>(stream_, n_, xf_, 1.0f);
    auto launch = [&]() {
        if (precision == LINA_PREC_FP64)
            dev::spmv<double, double>(stream_, A_, x_, nullptr, y_, dev::MODE_OUTER, nullptr, nullptr);
        else
            dev::spmv<float, float>(stream_, Af_, xf_, nullptr, yf_, dev::MODE_OUTER, nullptr, nullptr);
    };
    for (int i = 0; i < 3; i++) launch();
    hip_check(hipEventRecord(ev0_, stream_), "event");
    for (int i = 0; i < reps; i++) launch();
    hip_check(hipEventRecord(ev1_, stream_), "event");
    hip_check(hipEventSynchronize(ev1_), "event sync");
    float ms = 0;
    hip_check(hipEventElapsedTime(&ms, ev0_, ev1_), "elapsed");
    return ms / reps;
}

double DevMatrix::bytes(int precision) const {
    return precision == LINA_PREC_FP64 ? mat_bytes(A_) : mat_bytes(Af_);
}

template void Solver::smooth<double>(int, const double*, double*, int, bool, int);
template void Solver::smooth<float>(int, const float*, float*, int, bool, int);
template void Solver::vcycle<double>(int, const double*, double*, int);
template void Solver::vcycle<float>(int, const float*, float*, int);
template void Solver::fmg<double>(const double*, double*);
template void Solver::fmg<float>(const float*, float*);

}  // namespace lina
