// Host-side AMG setup of liblina_hip: C/F splitting, interpolation, Galerkin
// products and the coarsest-level LU, built so that every label, pattern and
// value equals what the reference's AMG_setup produces (LinA/Multigrid.c:
// 1613-1754) for the same input matrix.
//
// The reference stores matrices as row-lists and grows them with
// insert_sparse (binary search + realloc + memmove, linear_algebra.c:
// 1124-1159).  Here rows are CSR and every "insert" becomes an update of a
// dense per-row accumulator, applied in the same order as the reference
// applies its inserts, so the floating-point operations (and therefore the
// bits) are the same while the cost is linear in the work.
//
// Compiled with -ffp-contract=off: a*b+c is never fused, as in the reference.

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdlib>
#include <cstring>
#include <map>
#include <stdexcept>

#include <omp.h>

#include "lina_internal.hpp"

namespace lina {

// ------------------------------------------------------------- helpers

// get_diag (linear_algebra.c:3434-3439): value of (i,i) or 0 if absent
double get_diag(const Csr& A, int i) {
    const int* b = A.ci.data() + A.rp[i];
    const int* e = A.ci.data() + A.rp[i + 1];
    const int* p = std::lower_bound(b, e, i);
    if (p != e && *p == i) return A.v[A.rp[i] + (p - b)];
    return 0.0;
}

// Dense row accumulator with the insert_sparse/remove_element semantics:
// present[] marks stored entries, val[] holds their value, touched lists
// the columns to reset.
struct RowAcc {
    std::vector<double> val;
    std::vector<char> present;
    std::vector<int> touched;
    explicit RowAcc(int n = 0) : val(n, 0.0), present(n, 0) {}
    void resize(int n) {
        val.assign(n, 0.0);
        present.assign(n, 0);
        touched.clear();
    }
    // insert_sparse: new entry takes the value, existing entry accumulates
    inline void add(int j, double a) {
        if (present[j]) {
            val[j] += a;
        } else {
            present[j] = 1;
            val[j] = a;
            touched.push_back(j);
        }
    }
    // sparse_zero seeds (i,0.0): an existing entry with value +0.0
    inline void seed_zero(int j) {
        if (!present[j]) {
            present[j] = 1;
            val[j] = 0.0;
            touched.push_back(j);
        }
    }
    inline void remove(int j) {
        if (present[j]) {
            present[j] = 0;
            val[j] = 0.0;
        }
    }
    // sorted live entries; resets the accumulator
    void drain(std::vector<int>& cols, std::vector<double>& vals) {
        cols.clear();
        vals.clear();
        std::sort(touched.begin(), touched.end());
        touched.erase(std::unique(touched.begin(), touched.end()), touched.end());
        for (int j : touched) {
            if (present[j]) {
                cols.push_back(j);
                vals.push_back(val[j]);
            }
            present[j] = 0;
            val[j] = 0.0;
        }
        touched.clear();
    }
};

// delete_zeros row rule (linear_algebra.c:793-805) applied to a row that
// started from sparse_zero's (row,0.0) seed: exact zeros are dropped; an
// all-zero row collapses to the single entry (row, 0.0).
static void delete_zeros_row(int row, std::vector<int>& cols, std::vector<double>& vals) {
    if (cols.size() <= 1) return;
    size_t w = 0;
    for (size_t k = 0; k < cols.size(); k++) {
        if (vals[k] != 0.0) {
            cols[w] = cols[k];
            vals[w] = vals[k];
            w++;
        }
    }
    if (w == 0) {
        cols.assign(1, row);
        vals.assign(1, 0.0);
    } else {
        cols.resize(w);
        vals.resize(w);
    }
}

// ------------------------------------------- products (sparse_product)

// C = A*B with the reference's traversal: for each A(i,k)!=0 in row order,
// for each B(k,j)!=0 in row order, C(i,j) += A(i,k)*B(k,j); row i of C is
// seeded with (i,0.0); then delete_zeros (linear_algebra.c:186-222).
// Rows are independent: chunks of rows are computed in parallel (one
// accumulator per thread) and concatenated in row order.
template <typename RowFn>
static void rowwise(int n, int width, Csr& C, RowFn fn) {
    const int nchunk = std::max(1, std::min(n, omp_get_max_threads() * 8));
    std::vector<std::vector<int>> cc(nchunk), cl(nchunk);
    std::vector<std::vector<double>> cv(nchunk);
#pragma omp parallel
    {
        RowAcc acc(width);
        std::vector<int> cols;
        std::vector<double> vals;
#pragma omp for schedule(dynamic, 1)
        for (int c = 0; c < nchunk; c++) {
            int r0 = (int)((long long)n * c / nchunk), r1 = (int)((long long)n * (c + 1) / nchunk);
            for (int i = r0; i < r1; i++) {
                fn(i, acc, cols, vals);
                cc[c].insert(cc[c].end(), cols.begin(), cols.end());
                cv[c].insert(cv[c].end(), vals.begin(), vals.end());
                cl[c].push_back((int)cols.size());
            }
        }
    }
    C.rp.assign(n + 1, 0);
    size_t tot = 0;
    for (int c = 0; c < nchunk; c++) tot += cc[c].size();
    C.ci.clear();
    C.v.clear();
    C.ci.reserve(tot);
    C.v.reserve(tot);
    int i = 0;
    for (int c = 0; c < nchunk; c++) {
        C.ci.insert(C.ci.end(), cc[c].begin(), cc[c].end());
        C.v.insert(C.v.end(), cv[c].begin(), cv[c].end());
        for (int len : cl[c]) {
            C.rp[i + 1] = C.rp[i] + len;
            i++;
        }
        std::vector<int>().swap(cc[c]);
        std::vector<double>().swap(cv[c]);
    }
}

void sparse_product(const Csr& A, const Csr& B, int b_ncols, Csr& C) {
    const int n = A.n;
    C.n = n;
    C.ncols = b_ncols;
    rowwise(n, std::max(b_ncols, n), C, [&](int i, RowAcc& acc, std::vector<int>& cols, std::vector<double>& vals) {
        acc.seed_zero(i);
        for (int p = A.rp[i]; p < A.rp[i + 1]; p++) {
            double a = A.v[p];
            if (a == 0.0) continue;
            int k = A.ci[p];
            for (int q = B.rp[k]; q < B.rp[k + 1]; q++) {
                double b = B.v[q];
                if (b != 0.0) acc.add(B.ci[q], a * b);
            }
        }
        acc.drain(cols, vals);
        delete_zeros_row(i, cols, vals);
    });
}

// get_restricted_matrix (Multigrid.c:475-484): R*(A*P) + square_matrix_clean
void galerkin(const Csr& A, const Csr& R, const Csr& P, Csr& Ac) {
    Csr AP;
    sparse_product(A, P, P.ncols, AP);
    sparse_product(R, AP, P.ncols, Ac);
    // square_matrix_clean (linear_algebra.c:3100-3120)
    const int n = Ac.n;
    Csr out;
    out.n = n;
    out.ncols = n;
    out.rp.assign(n + 1, 0);
    for (int i = 0; i < n; i++) {
        for (int p = Ac.rp[i]; p < Ac.rp[i + 1]; p++) {
            int j = Ac.ci[p];
            double v = Ac.v[p];
            if (j >= n) {
                if (v == 0.0) continue;
                throw std::runtime_error("erroneous coarse matrix (column outside the square part)");
            }
            if (v == 0.0 && j != i) continue;
            out.ci.push_back(j);
            out.v.push_back(v);
        }
        out.rp[i + 1] = (int)out.ci.size();
    }
    Ac = std::move(out);
}

// ------------------------------------------------- strength of connection

struct Lists {  // ragged int lists, CSR-like
    std::vector<int> ptr, idx;
    int size(int i) const { return ptr[i + 1] - ptr[i]; }
    const int* row(int i) const { return idx.data() + ptr[i]; }
};

// get_max_coupled (Multigrid.c:86-97): position of the first most negative
// off-diagonal value, -1 if there is none
static int max_coupled(const Csr& A, int i) {
    double mn = 0.0;
    int index = -1;
    for (int p = A.rp[i]; p < A.rp[i + 1]; p++) {
        if (A.ci[p] != i && A.v[p] < mn) {
            mn = A.v[p];
            index = p;
        }
    }
    return index;
}

// get_strong_connection_sets (Multigrid.c:156-184, 240-258)
static void strong_sets(const Csr& A, const int* levels, double eps_minus, double eps_plus, Lists& S, Lists& T) {
    const int n = A.n;
    S.ptr.assign(n + 1, 0);
    S.idx.clear();
    for (int i = 0; i < n; i++) {
        int mi = max_coupled(A, i);
        if (mi >= 0) {
            double mn = -eps_minus * std::fabs(A.v[mi]);
            double mx = eps_plus * std::fabs(A.v[mi]);
            for (int p = A.rp[i]; p < A.rp[i + 1]; p++) {
                int J = A.ci[p];
                if (J == i) continue;
                double a = (levels[i] == levels[J]) ? A.v[p] : 0.0;
                if (a < mn || a > mx) S.idx.push_back(J);
            }
        }
        S.ptr[i + 1] = (int)S.idx.size();
    }
    // transpose: T[k] lists j (ascending) with k in S[j]
    T.ptr.assign(n + 1, 0);
    for (int k : S.idx) T.ptr[k + 1]++;
    for (int i = 0; i < n; i++) T.ptr[i + 1] += T.ptr[i];
    T.idx.assign(S.idx.size(), 0);
    std::vector<int> fill(T.ptr.begin(), T.ptr.end() - 1);
    for (int j = 0; j < n; j++)
        for (int p = S.ptr[j]; p < S.ptr[j + 1]; p++) T.idx[fill[S.idx[p]]++] = j;
}

// ------------------------------------------------------------ coarsener

// AMG_get_lambda (Multigrid.c:364-375)
static inline double get_lambda(const int* label, const Lists& T, int i, int size) {
    int undec = 0, fine = 0;
    const int* t = T.row(i);
    for (int p = 0; p < T.size(i); p++) {
        int J = t[p];
        if (label[J] == UNDECIDED) undec++;
        if (label[J] == FINE) fine++;
    }
    return (double)i / size + undec + 2 * fine;
}

// The reference keeps candidates in a red-black tree keyed by Importance
// (red_black_tree.c): equal keys are inserted to the right of existing ones
// (TreeInsertHelp, :174-202) and RBLargest takes the rightmost node
// (:513-523).  std::multimap inserts at the upper bound of the equal range and
// prev(end()) is the rightmost element, which is the same order.
using Tree = std::multimap<double, int>;

struct CoarsenState {
    Tree tree;
    std::vector<Tree::iterator> node;
    std::vector<char> in_tree;
    explicit CoarsenState(int n) : node(n), in_tree(n, 0) {}
    void insert(int I, double key) {
        node[I] = tree.insert(std::make_pair(key, I));
        in_tree[I] = 1;
    }
    void erase(int I) {
        tree.erase(node[I]);
        in_tree[I] = 0;
    }
};

// AMG_coarsener (Multigrid.c:377-473).  `levels` is indexed with the
// matrix's own row indices exactly as the reference does, including the
// aggressive second pass where it receives the fine-level array.
static void coarsener(const Csr& A, int* label, double* imp, const Lists& S, const Lists& T,
                      CoarsenState& st, const int* levels, int size, int seed) {
    int level = levels[seed];
    int next = seed;
    int cursor = 0;  // rows below cursor are decided or of another level
    do {
        label[next] = COARSE;
        for (int p = A.rp[next]; p < A.rp[next + 1]; p++) {
            int I = A.ci[p];
            if (label[I] == UNDECIDED && A.v[p] != 0.0 && !st.in_tree[I]) st.insert(I, imp[I]);
        }
        {
            const int* t = T.row(next);
            for (int q = 0; q < T.size(next); q++) {
                int I = t[q];
                if (st.in_tree[I]) {
                    label[I] = FINE;
                    st.erase(I);
                }
            }
        }
        for (int p = A.rp[next]; p < A.rp[next + 1]; p++) {
            int I = A.ci[p];
            if (A.v[p] != 0.0 && st.in_tree[I]) {
                double lambda = get_lambda(label, T, I, size);
                if (lambda != imp[I]) {
                    st.erase(I);
                    imp[I] = lambda;
                    st.insert(I, imp[I]);
                }
            }
        }
        for (int p = A.rp[next]; p < A.rp[next + 1]; p++) {
            int I = A.ci[p];
            if (I == next) continue;
            for (int q = A.rp[I]; q < A.rp[I + 1]; q++) {
                int J = A.ci[q];
                if (label[J] == UNDECIDED && A.v[q] != 0.0 && !st.in_tree[J]) st.insert(J, imp[J]);
            }
            const int* s = S.row(I);
            for (int q = 0; q < S.size(I); q++) {
                int J = s[q];
                if (label[J] == UNDECIDED) {
                    double lambda = get_lambda(label, T, J, size);
                    if (lambda != imp[J]) {
                        if (st.in_tree[J]) st.erase(J);
                        imp[J] = lambda;
                        st.insert(J, imp[J]);
                    }
                }
            }
        }
        if (!st.tree.empty()) {
            auto last = std::prev(st.tree.end());
            next = last->second;
            st.erase(next);
        } else {
            next = -1;
            for (int i = cursor; i < size; i++)
                if (label[i] == UNDECIDED && levels[i] == level) {
                    next = i;
                    cursor = i;
                    break;
                }
            if (next < 0) cursor = size;
        }
    } while (next >= 0);
}

// AMG_is_isolated (Multigrid.c:498-503): exactly one nonzero value
static bool is_isolated(const Csr& A, int i) {
    int nz = 0;
    for (int p = A.rp[i]; p < A.rp[i + 1]; p++)
        if (A.v[p] != 0.0) nz++;
    return nz == 1;
}

enum { STAND_ALONE = 0, PREMODE = 1 };

// standard_coarsening (Multigrid.c:506-548)
static int standard_coarsening(const Csr& A, int* label, Lists& S, Lists& T, const int* levels, int dof,
                               int mode, const SetupParams& prm) {
    const int n = A.n;
    strong_sets(A, levels, prm.eps_minus, prm.eps_plus, S, T);
    std::vector<double> imp(n, 0.0);
    for (int i = 0; i < n; i++) {
        if (S.size(i) == 0)
            label[i] = is_isolated(A, i) ? ISOLATED : COARSE;
        else
            label[i] = UNDECIDED;
    }
    CoarsenState st(n);
    for (int d = 0; d < dof; d++) {
        int seed = -1;
        for (int j = 0; j < n; j++)
            if (label[j] == UNDECIDED && levels[j] == d) {
                seed = j;
                break;
            }
        if (seed >= 0) coarsener(A, label, imp.data(), S, T, st, levels, n, seed);
    }
    int coarse = 0;
    for (int i = 0; i < n; i++) {
        if (label[i] == UNDECIDED) label[i] = FINE;
        if (label[i] == COARSE) coarse++;
    }
    if (mode == STAND_ALONE)
        for (int i = 0; i < n; i++)
            if (label[i] == ISOLATED) coarse++;
    return coarse;
}

// get_2nd_max_coupled (Multigrid.c:99-119)
static double second_max_coupled(const Csr& A, int i) {
    double mx = 0.0;
    for (int p = A.rp[i]; p < A.rp[i + 1]; p++) {
        double val = A.v[p];
        int J = A.ci[p];
        double diag = get_diag(A, J);
        if (J != i && val != 0.0)
            for (int q = A.rp[J]; q < A.rp[J + 1]; q++) {
                double val2 = A.v[q];
                double pp = std::fabs(val * val2 / diag);
                int K = A.ci[q];
                if (K != J && K != i && val2 != 0.0 && pp > mx) mx = pp;
            }
    }
    return mx;
}

// aggressive_coarsening (Multigrid.c:550-607) with the distance-2 strength of
// Get_2nd_strong_connected (:187-217) and get_2nd_strong_connection_sets
// (:260-292)
static int aggressive_coarsening(const Csr& A, int* label, Lists& S, Lists& T, const int* levels, int dof,
                                 const SetupParams& prm, std::vector<std::string>& warn) {
    const int n = A.n;
    std::vector<int> pre(n);
    int pre_size = standard_coarsening(A, pre.data(), S, T, levels, dof, PREMODE, prm);
    for (int i = 0; i < n; i++) label[i] = pre[i];
    std::vector<int> map(n, -1), inv(pre_size);
    {
        int k = 0;
        for (int i = 0; i < n; i++)
            if (pre[i] == COARSE) {
                map[i] = k;
                inv[k] = i;
                k++;
            }
    }
    // Pre_A: pattern of distance-2 coarse neighbours, values = multiplicity,
    // diagonal seeded with 0.0 (sparse_zero).  Matrix_2nd / Transpose_2nd keep
    // duplicates exactly as the reference's realloc'd lists do.
    Csr PA;
    PA.n = pre_size;
    PA.ncols = pre_size;
    PA.rp.assign(pre_size + 1, 0);
    Lists M2;
    M2.ptr.assign(pre_size + 1, 0);
    RowAcc acc(pre_size);
    std::vector<int> cols;
    std::vector<double> vals;
    for (int i = 0; i < n; i++) {
        if (pre[i] != COARSE) continue;
        int I = map[i];
        double max_val = second_max_coupled(A, i);
        if (max_val > 0) {
            double mn = -prm.eps_minus * max_val;
            double mx = prm.eps_plus * max_val;
            for (int p = A.rp[i]; p < A.rp[i + 1]; p++) {
                int J = A.ci[p];
                double a = A.v[p];
                double d = get_diag(A, J);
                if (J != i && a != 0.0 && levels[i] == levels[J]) {
                    for (int q = A.rp[J]; q < A.rp[J + 1]; q++) {
                        int K = A.ci[q];
                        double b = A.v[q];
                        if (pre[K] == COARSE && K != J && K != i && a * b / d > 0 && levels[K] == levels[i]) {
                            if (-a * b / d < mn || a * b / d > mx) M2.idx.push_back(map[K]);
                        }
                    }
                }
            }
        }
        M2.ptr[I + 1] = (int)M2.idx.size();
        acc.seed_zero(I);
        for (int p = A.rp[i]; p < A.rp[i + 1]; p++) {
            int J = A.ci[p];
            if (J != i && A.v[p] != 0.0)
                for (int q = A.rp[J]; q < A.rp[J + 1]; q++) {
                    int K = A.ci[q];
                    if (pre[K] == COARSE && K != J && K != i && A.v[q] != 0.0) acc.add(map[K], 1.0);
                }
        }
        acc.drain(cols, vals);
        PA.ci.insert(PA.ci.end(), cols.begin(), cols.end());
        PA.v.insert(PA.v.end(), vals.begin(), vals.end());
        PA.rp[I + 1] = (int)PA.ci.size();
    }
    Lists T2;
    T2.ptr.assign(pre_size + 1, 0);
    for (int k : M2.idx) T2.ptr[k + 1]++;
    for (int I = 0; I < pre_size; I++) T2.ptr[I + 1] += T2.ptr[I];
    T2.idx.assign(M2.idx.size(), 0);
    {
        std::vector<int> fill(T2.ptr.begin(), T2.ptr.end() - 1);
        for (int I = 0; I < pre_size; I++)
            for (int p = M2.ptr[I]; p < M2.ptr[I + 1]; p++) T2.idx[fill[M2.idx[p]]++] = I;
    }
    // Importance_2nd is malloc'd without initialisation in the reference
    // (Multigrid.c:566); it is taken as zero-filled here.
    std::vector<int> label2(pre_size, UNDECIDED);
    std::vector<double> imp2(pre_size, 0.0);
    CoarsenState st(pre_size);
    for (int d = 0; d < dof; d++) {
        int seed = -1;
        for (int I = 0; I < pre_size; I++)
            if (levels[inv[I]] == d) {
                seed = I;
                break;
            }
        if (seed < 0) {
            warn.push_back("aggressive coarsening: no pre-coarse point of dof component, pass skipped");
            continue;
        }
        // quirk kept: the second pass indexes the fine-level Levels[] array
        // with pre-coarse indices (Multigrid.c:389,462 called from 580-581)
        coarsener(PA, label2.data(), imp2.data(), M2, T2, st, levels, pre_size, seed);
    }
    int coarse = 0;
    for (int i = 0; i < n; i++) {
        if (label[i] == COARSE) {
            label[i] = label2[map[i]];
            if (label[i] == COARSE) coarse++;
        } else if (label[i] == ISOLATED) {
            coarse++;
        }
    }
    return coarse;
}

// ------------------------------------------------------- interpolation

// AMG_simple_standard_interpolation (Multigrid.c:989-1079) followed by
// sparse_approximate (linear_algebra.c:1418-1442) and normalize_row_sum
// (:455-464), as set_prolongation_map does (Multigrid.c:1265-1286).
static void interpolation(const Csr& A, const int* label, const int* levels, const std::vector<int>& map,
                          int coarse_size, int passes, double tol, Csr& P, std::vector<std::string>& warn) {
    const int n = A.n;
    // D0: diagonal used to build and eliminate Q (first loop of the reference);
    // the second loop overwrites D[i] of a FINE row only for that row's weights.
    std::vector<double> D0(n);
    for (int i = 0; i < n; i++) {
        D0[i] = get_diag(A, i);
        if (D0[i] == 0.0) throw std::runtime_error("zero diagonal: standard interpolation not possible");
    }
    // rows with no coarse interpolation node: the reference `break`s there and
    // leaves this and every later row of P at sparse_zero's (i,0.0)
    std::vector<char> stops(n, 0);
    P.n = n;
    P.ncols = coarse_size;
    rowwise(n, n, P, [&](int i, RowAcc& acc, std::vector<int>& pc, std::vector<double>& pv) {
        thread_local std::vector<int> snap_c, rc;
        thread_local std::vector<double> snap_v, rv;
        pc.clear();
        pv.clear();
        if (label[i] == FINE) {
            // Q row i = A row i / D0[i], then `passes` elimination passes over a
            // snapshot of the row (clone_list/clone_vector)
            for (int p = A.rp[i]; p < A.rp[i + 1]; p++) acc.add(A.ci[p], A.v[p] / D0[i]);
            for (int k = 0; k < passes; k++) {
                acc.drain(snap_c, snap_v);
                for (size_t t = 0; t < snap_c.size(); t++) acc.add(snap_c[t], snap_v[t]);
                for (size_t t = 0; t < snap_c.size(); t++) {
                    int J = snap_c[t];
                    double q = snap_v[t];
                    if (J != i && label[J] == FINE && q != 0.0) {
                        acc.remove(J);
                        for (int l = A.rp[J]; l < A.rp[J + 1]; l++) {
                            int L = A.ci[l];
                            double a = A.v[l] / D0[J];
                            if (L != J && a != 0.0) acc.add(L, -a * q);
                        }
                    }
                }
            }
            acc.drain(rc, rv);
            // drop couplings to other dof components (order preserving)
            size_t w = 0;
            for (size_t t = 0; t < rc.size(); t++)
                if (levels[rc[t]] == levels[i]) {
                    rc[w] = rc[t];
                    rv[w] = rv[t];
                    w++;
                }
            rc.resize(w);
            rv.resize(w);
            double s_p = 0, c_p = 0, s_m = 0, c_m = 0;
            double Di = D0[i];
            for (size_t t = 0; t < rc.size(); t++) {
                int J = rc[t];
                double q = rv[t];
                if (J != i) {
                    if (q < 0) {
                        s_m += q;
                        if (label[J] == COARSE) c_m += q;
                    } else {
                        s_p += q;
                        if (label[J] == COARSE) c_p += q;
                    }
                } else {
                    Di = q;
                }
            }
            double alpha = 0, beta = 0;
            if (c_m != 0) alpha = -s_m / c_m; else Di += s_m;
            if (c_p != 0) beta = -s_p / c_p; else Di += s_p;
            if (alpha == 0 && beta == 0) {
                stops[i] = 1;
                pc.push_back(i);
                pv.push_back(0.0);
                return;
            }
            alpha /= Di;
            beta /= Di;
            // P row: sparse_zero seed (i,0.0) merged with the inserts
            acc.seed_zero(i);
            for (size_t t = 0; t < rc.size(); t++) {
                int J = rc[t];
                double q = rv[t];
                if (J != i && label[J] == COARSE) {
                    if (q < 0 && alpha != 0) acc.add(map[J], alpha * q);
                    if (q > 0 && beta != 0) acc.add(map[J], beta * q);
                }
            }
            acc.drain(pc, pv);
        } else {
            acc.seed_zero(i);
            acc.add(map[i], 1.0);
            acc.drain(pc, pv);
        }
        delete_zeros_row(i, pc, pv);
        // sparse_approximate: the "diagonal" test compares the coarse column
        // with the fine row index, as in the reference
        double mx = 0.0;
        for (size_t t = 0; t < pc.size(); t++) {
            double a = std::fabs(pv[t]);
            if (pc[t] != i && a > mx) mx = a;
        }
        size_t w = 0;
        for (size_t t = 0; t < pc.size(); t++) {
            if (pc[t] != i && std::fabs(pv[t]) < mx * tol) continue;
            pc[w] = pc[t];
            pv[w] = pv[t];
            w++;
        }
        pc.resize(w);
        pv.resize(w);
        // normalize_row_sum with Row_Sums == NULL
        double sum = 0;
        for (size_t t = 0; t < pc.size(); t++) sum += pv[t];
        if (sum != 0)
            for (size_t t = 0; t < pc.size(); t++) pv[t] *= 1. / sum;
    });
    int first = -1;
    for (int i = 0; i < n; i++)
        if (stops[i]) {
            first = i;
            break;
        }
    if (first >= 0) {
        warn.push_back("interpolation: no coarse interpolation nodes at row " + std::to_string(first) +
                       " (reference stops filling P there)");
        Csr Q;
        Q.n = n;
        Q.ncols = coarse_size;
        Q.rp.assign(n + 1, 0);
        for (int i = 0; i < n; i++) {
            if (i < first) {
                for (int p = P.rp[i]; p < P.rp[i + 1]; p++) {
                    Q.ci.push_back(P.ci[p]);
                    Q.v.push_back(P.v[p]);
                }
            } else {
                Q.ci.push_back(i);
                Q.v.push_back(0.0);
            }
            Q.rp[i + 1] = (int)Q.ci.size();
        }
        P = std::move(Q);
    }
}

// get_AMG_transposed (Multigrid.c:308-321): R = P^T without P's zeros; row k
// of R carries sparse_zero's (k,0.0) seed unless a value lands on column k
static void transpose_restriction(const Csr& P, int coarse_size, Csr& R) {
    R.n = coarse_size;
    R.ncols = P.n;
    R.rp.assign(coarse_size + 1, 0);
    std::vector<int> cnt(coarse_size, 0);
    std::vector<char> seed_hit(coarse_size, 0);
    for (int i = 0; i < P.n; i++)
        for (int p = P.rp[i]; p < P.rp[i + 1]; p++)
            if (P.v[p] != 0.0) {
                int k = P.ci[p];
                cnt[k]++;
                if (k == i) seed_hit[k] = 1;
            }
    for (int k = 0; k < coarse_size; k++) R.rp[k + 1] = R.rp[k] + cnt[k] + (seed_hit[k] ? 0 : 1);
    R.ci.assign(R.rp[coarse_size], 0);
    R.v.assign(R.rp[coarse_size], 0.0);
    std::vector<int> fill(R.rp.begin(), R.rp.end() - 1);
    // entries arrive with increasing fine index i; an unhit seed (k,0.0) is
    // placed before the first entry with i > k, or last
    std::vector<char> seed_done(coarse_size, 0);
    for (int i = 0; i < P.n; i++)
        for (int p = P.rp[i]; p < P.rp[i + 1]; p++) {
            double a = P.v[p];
            if (a == 0.0) continue;
            int k = P.ci[p];
            if (!seed_hit[k] && !seed_done[k] && k < i) {
                R.ci[fill[k]] = k;
                R.v[fill[k]] = 0.0;
                fill[k]++;
                seed_done[k] = 1;
            }
            R.ci[fill[k]] = i;
            R.v[fill[k]] = a;  // 0.0 + a for a hit seed: same value
            fill[k]++;
        }
    for (int k = 0; k < coarse_size; k++)
        if (!seed_hit[k] && !seed_done[k]) {
            R.ci[fill[k]] = k;
            R.v[fill[k]] = 0.0;
            fill[k]++;
        }
}

// ------------------------------------------------------ coarsest level

// sparse_LU_factorization (linear_algebra.c:3040-3062) splits the coarsest
// matrix into a dense system over the rows with more than one nonzero value
// (rows and columns restricted to those indices, sub_convert_sparse_to_array
// :2993-3023) and trivial rows solved as x_i = f_i / a_ii (L = 1, U = diag,
// :3034-3037).  The dense system is extracted here; its inverse is formed on
// the device (Gauss-Jordan, no pivoting, as LU_factorzation does not pivot).
void coarse_lu(const Csr& A0, CoarseLU& lu) {
    const int n = A0.n;
    lu = CoarseLU();
    lu.n = n;
    std::vector<int> map(n, -1);
    for (int i = 0; i < n; i++) {
        int nz = 0;
        for (int p = A0.rp[i]; p < A0.rp[i + 1]; p++)
            if (A0.v[p] != 0.0) nz++;
        if (nz > 1) {
            map[i] = (int)lu.dense_idx.size();
            lu.dense_idx.push_back(i);
        } else {
            lu.trivial_idx.push_back(i);
            lu.trivial_diag.push_back(get_diag(A0, i));
        }
    }
    const int N = (int)lu.dense_idx.size();
    lu.dense.assign((size_t)N * N, 0.0);
    for (int I = 0; I < N; I++) {
        int i = lu.dense_idx[I];
        for (int p = A0.rp[i]; p < A0.rp[i + 1]; p++) {
            int J = map[A0.ci[p]];
            if (J >= 0) lu.dense[(size_t)I * N + J] = A0.v[p];
        }
    }
}

// ------------------------------------------------------------ colouring

// Greedy colouring in natural row order over the symmetrised pattern of
// nonzero off-diagonal values: colour(i) = smallest colour not used by an
// already coloured neighbour.
std::vector<int> greedy_colouring(const Csr& A, int* ncolours) {
    const int n = A.n;
    std::vector<int> tptr(n + 1, 0);
    for (int i = 0; i < n; i++)
        for (int p = A.rp[i]; p < A.rp[i + 1]; p++)
            if (A.ci[p] != i && A.v[p] != 0.0) tptr[A.ci[p] + 1]++;
    for (int i = 0; i < n; i++) tptr[i + 1] += tptr[i];
    std::vector<int> tadj(tptr[n]);
    {
        std::vector<int> fill(tptr.begin(), tptr.end() - 1);
        for (int i = 0; i < n; i++)
            for (int p = A.rp[i]; p < A.rp[i + 1]; p++)
                if (A.ci[p] != i && A.v[p] != 0.0) tadj[fill[A.ci[p]]++] = i;
    }
    std::vector<int> color(n, 0), mark(64, -1);
    int ncol = 0;
    for (int i = 0; i < n; i++) {
        for (int p = A.rp[i]; p < A.rp[i + 1]; p++) {
            int k = A.ci[p];
            if (k < i && A.v[p] != 0.0) mark[color[k]] = i;
        }
        for (int p = tptr[i]; p < tptr[i + 1]; p++) {
            int k = tadj[p];
            if (k < i) mark[color[k]] = i;
        }
        int c = 0;
        while (c < (int)mark.size() && mark[c] == i) c++;
        if (c == (int)mark.size()) mark.resize(mark.size() * 2, -1);
        color[i] = c;
        ncol = std::max(ncol, c + 1);
    }
    *ncolours = ncol;
    return color;
}

// ------------------------------------------------------------ ILU(0)

// incomplete_LU_factorization (linear_algebra.c:1952-1982) with
// row_times_row (:1795-1819).  L holds the lower part with its diagonal, U
// the unit upper part, both rows sorted.  Returns 0 on a zero pivot.
int ilu0(const Csr& A, Csr& L, Csr& U) {
    const int n = A.n;
    // U is built column-wise first (Ucol[j] = entries U(i<=j, j) by row i),
    // exactly as the reference inserts U(j,i) and transposes at the end.
    std::vector<std::vector<int>> Lc(n), Uc(n);
    std::vector<std::vector<double>> Lv(n), Uv(n);
    auto rtr = [&](int ia, int ib, int jmax) {
        // row_times_row(L, Ucol, ia, ib, 0, jmax): sum over common columns <= jmax
        double sum = 0.0;
        const auto& ac = Lc[ia];
        const auto& av = Lv[ia];
        const auto& bc = Uc[ib];
        const auto& bv = Uv[ib];
        size_t s = 0;
        for (size_t l = 0; l < ac.size(); l++) {
            int k = ac[l];
            if (k > jmax) break;
            while (s < bc.size() && bc[s] < k) s++;
            if (s < bc.size() && bc[s] == k) sum += av[l] * bv[s];
        }
        return sum;
    };
    auto ins = [](std::vector<int>& c, std::vector<double>& v, int j, double a) {
        auto it = std::lower_bound(c.begin(), c.end(), j);
        size_t pos = it - c.begin();
        if (it != c.end() && *it == j) v[pos] += a;
        else {
            c.insert(it, j);
            v.insert(v.begin() + pos, a);
        }
    };
    for (int i = 0; i < n; i++) {
        Lc[i].push_back(i); Lv[i].push_back(0.0);    // sparse_zero seeds
    }
    for (int i = 0; i < n; i++) {
        Uc[i].push_back(i); Uv[i].push_back(0.0);
    }
    for (int i = 0; i < n; i++) {
        ins(Uc[i], Uv[i], i, 1.0);
        for (int p = A.rp[i]; p < A.rp[i + 1]; p++) {
            int j = A.ci[p];
            if (j <= i) {
                double sum = A.v[p];
                sum -= rtr(i, j, j - 1);
                ins(Lc[i], Lv[i], j, sum);
            } else {
                double sum = A.v[p];
                sum -= rtr(i, j, i - 1);
                double d = 0.0;
                {
                    auto it = std::lower_bound(Lc[i].begin(), Lc[i].end(), i);
                    if (it != Lc[i].end() && *it == i) d = Lv[i][it - Lc[i].begin()];
                }
                if (d != 0) sum /= d;
                else return 0;
                ins(Uc[j], Uv[j], i, sum);
            }
        }
    }
    L.n = n; L.ncols = n; L.rp.assign(n + 1, 0); L.ci.clear(); L.v.clear();
    for (int i = 0; i < n; i++) {
        L.ci.insert(L.ci.end(), Lc[i].begin(), Lc[i].end());
        L.v.insert(L.v.end(), Lv[i].begin(), Lv[i].end());
        L.rp[i + 1] = (int)L.ci.size();
    }
    // get_transpose of Ucol (seeded with sparse_zero rows (k,0.0))
    std::vector<std::vector<int>> Tc(n);
    std::vector<std::vector<double>> Tv(n);
    for (int k = 0; k < n; k++) { Tc[k].push_back(k); Tv[k].push_back(0.0); }
    for (int i = 0; i < n; i++)
        for (size_t r = 0; r < Uc[i].size(); r++) ins(Tc[Uc[i][r]], Tv[Uc[i][r]], i, Uv[i][r]);
    U.n = n; U.ncols = n; U.rp.assign(n + 1, 0); U.ci.clear(); U.v.clear();
    for (int i = 0; i < n; i++) {
        U.ci.insert(U.ci.end(), Tc[i].begin(), Tc[i].end());
        U.v.insert(U.v.end(), Tv[i].begin(), Tv[i].end());
        U.rp[i + 1] = (int)U.ci.size();
    }
    return 1;
}

// ------------------------------------------------------------ AMG_setup

static bool verbose() {
    static const bool v = getenv("LINA_VERBOSE") != nullptr;  // thread-safe initialisation
    return v;
}

static double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

void amg_setup(const Csr& Ain, const SetupParams& prm, HostHierarchy& H) {
    const int depth = prm.depth;
    double t0 = now_s();
    if (depth < 2) throw std::runtime_error("AMG depth must be >= 2");
    H = HostHierarchy();
    H.prm = prm;
    H.lv.resize(depth);
    const int n = Ain.n;
    // clone + make_positive_diag (linear_algebra.c:2612-2620)
    Csr A = Ain;
    H.signs.assign(n, 1.0);
    for (int i = 0; i < n; i++)
        if (get_diag(A, i) < 0) {
            for (int p = A.rp[i]; p < A.rp[i + 1]; p++) A.v[p] *= -1.;
            H.signs[i] = -1.0;
        }
    H.lv[depth - 1].A = std::move(A);
    H.lv[depth - 1].fine = n;
    // AMG_init_finest_tupels (Multigrid.c:785-804): blocked dof numbering
    std::vector<int> levels(n, 0);
    {
        int mesh = n / prm.dof;
        for (int i = 0; i < mesh; i++)
            for (int j = 1; j <= prm.dof; j++) levels[(j - 1) * mesh + i] = j - 1;
    }
    for (int i = depth - 1; i > 0; i--) {
        HostLevel& L = H.lv[i];
        L.levels = levels;
        const Csr& Af = L.A;
        Lists S, T;
        L.labels.assign(L.fine, UNDECIDED);
        int coarse;
        if (prm.coarsening == AGGRESSIVE && i == depth - 1)
            coarse = aggressive_coarsening(Af, L.labels.data(), S, T, levels.data(), prm.dof, prm, H.warnings);
        else
            coarse = standard_coarsening(Af, L.labels.data(), S, T, levels.data(), prm.dof, STAND_ALONE, prm);
        L.coarse = coarse;
        // index map with isolated points (get_AMG_index_map, WITH_ISOLATED)
        std::vector<int> map(L.fine, -1);
        int k = 0;
        for (int r = 0; r < L.fine; r++)
            if (L.labels[r] != FINE) map[r] = k++;
        if (k != coarse)
            throw std::runtime_error("level " + std::to_string(i) +
                                     ": coarsening left undecided points (the reference would write past P)");
        // AMG_set_tupels: levels of the coarse points
        std::vector<int> new_levels(std::max(k, coarse), 0);
        for (int r = 0; r < L.fine; r++)
            if (map[r] >= 0 && map[r] < (int)new_levels.size()) new_levels[map[r]] = levels[r];
        int passes = (prm.coarsening == AGGRESSIVE && i == depth - 1) ? 2 : 1;
        interpolation(Af, L.labels.data(), levels.data(), map, coarse, passes, prm.matrix_tol, L.P, H.warnings);
        transpose_restriction(L.P, coarse, L.R);
        double t1 = now_s();
        HostLevel& C = H.lv[i - 1];
        galerkin(Af, L.R, L.P, C.A);
        C.fine = coarse;
        if (verbose()) {
            int iso = 0;
            for (int r = 0; r < L.fine; r++) iso += L.labels[r] == ISOLATED;
            fprintf(stderr, "[lina] level %d: %d -> %d (isolated %d) nnz A %d P %d Ac %d  split+P %.2fs RAP %.2fs\n", i,
                    L.fine, coarse, iso, Af.nnz(), L.P.nnz(), C.A.nnz(), t1 - t0, now_s() - t1);
        }
        t0 = now_s();
        levels.assign(new_levels.begin(), new_levels.begin() + coarse);
        if (i == depth - 1) {
            for (int r = 0; r < L.fine; r++)
                if (L.labels[r] == ISOLATED) {
                    H.iso_rows.push_back(r);
                    double a = 0.0;
                    for (int p = Af.rp[r]; p < Af.rp[r + 1]; p++)
                        if (Af.v[p] != 0.0) {
                            a = Af.v[p];
                            break;
                        }
                    H.iso_a.push_back(a);
                }
        }
    }
    H.lv[0].levels = levels;
    coarse_lu(H.lv[0].A, H.lu);
    if (verbose()) fprintf(stderr, "[lina] coarsest: %d rows, dense %zu\n", H.lv[0].A.n, H.lu.dense_idx.size());
}

// AMG_recompute_coarse_matrices (Multigrid.c:1756-1780): new finest values,
// P and R kept, Galerkin products and the coarsest LU rebuilt.  Signs are not
// recomputed (as in the reference); the finest matrix is taken as given.
void amg_recompute(HostHierarchy& H, const Csr& A_new) {
    const int depth = (int)H.lv.size();
    H.lv[depth - 1].A = A_new;
    for (int i = depth - 1; i > 0; i--) galerkin(H.lv[i].A, H.lv[i].R, H.lv[i].P, H.lv[i - 1].A);
    coarse_lu(H.lv[0].A, H.lu);
}

}  // namespace lina
