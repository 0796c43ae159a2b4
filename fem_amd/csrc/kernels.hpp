// Launch interface of the device kernels (kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include <cstdint>

namespace lina {
namespace dev {

// products of one row block staged in LDS: at most block_nnz<VT>() entries
// (2048 for both value types; 4096-entry FP32 blocks measured slower) and at
// most BLOCK_ROWS rows per block (one lane per row in the second phase)
#ifndef LINA_BLOCK_NNZ
#define LINA_BLOCK_NNZ 2048
#endif
#ifndef LINA_BLOCK_NNZ32
#define LINA_BLOCK_NNZ32 2048
#endif
#ifndef LINA_BLOCK_ROWS
#define LINA_BLOCK_ROWS 256
#endif
constexpr int BLOCK_NNZ = LINA_BLOCK_NNZ;
constexpr int BLOCK_NNZ32 = LINA_BLOCK_NNZ32;
constexpr int BLOCK_ROWS = LINA_BLOCK_ROWS;
template <typename VT>
constexpr int block_nnz() {
    return sizeof(VT) == 8 ? BLOCK_NNZ : BLOCK_NNZ32;
}

// MODE_OUTER is MODE_SET for the finest-level operator of the outer Krylov
// loop; a separate instantiation so that profiles report it on its own.
// MODE_ADDB: y = b + A x (a prolongation written beside its base vector);
// MODE_SETX: y = A x followed by the epilogue of an Epi (below).
enum { MODE_SET = 0, MODE_RES = 1, MODE_ADD = 2, MODE_OUTER = 3, MODE_ADDB = 4, MODE_SETX = 5 };
enum { JAC_REF = 0, JAC_L1 = 1 };

// device scalar slots used by the Krylov drivers
enum {
    S_S = 0,      // s = R.Z       (S_SOLD = S_S + 1 by k_finalize's shift)
    S_SOLD = 1,
    S_PQ = 2,     // P.Q
    S_R = 3,      // |R|
    S_H = 4,      // |B|
    S_R0 = 5,     // r0
    S_T0 = 6,     // scratch slots
    S_T1 = 7,
    S_T2 = 8,
    S_T3 = 9,
    S_T4 = 10,
    S_T5 = 11,
    S_AMG0 = 12,  // sum of squares of the preconditioner's right-hand side
    S_STOP = 13,  // pcg's device-side stop flag (nonzero: converged, NaN or out of iterations)
    S_IT = 14,    // pcg iterations completed on the device
    S_COUNT = 16
};

// device CSR with its LDS row-block schedule
template <typename VT>
struct DMat {
    int n = 0, ncols = 0, nnz = 0;
    int* rp = nullptr;
    int* ci = nullptr;
    VT* v = nullptr;
    int nblk = 0;
    int2* bk = nullptr;  // nblk+1 (first row, first entry) pairs of the row blocks
    // 16-bit column codes: a block's columns form at most two clusters, each
    // spanning < 32768, with base columns cb[block] = (b0, b1); an entry's
    // column is (code & 0x8000 ? b1 : b0) + (code & 0x7fff).  b0 < 0: the
    // block's columns are read from ci.  cs == nullptr: ci everywhere.
    uint16_t* cs = nullptr;
    int2* cb = nullptr;
    double idx_bytes = 0;  // column index bytes the row kernels read (host-side accounting)
    // batches: the first and last system of each row block's rows, and the
    // systems' active flags; a block whose systems are all inactive is skipped
    int2* bspec = nullptr;
    const double* active = nullptr;
    int lpr = 1;         // lanes reducing one row from LDS: 1 (sequential), 4, 8, 16, 32
    int* dpos = nullptr; // position of each row's diagonal entry in ci/v (-1: none)
    VT* l1 = nullptr;    // sum_j |a_ij| of each row (l1-Jacobi), summed in column order
};

// Epilogue of a MODE_SETX product y = A x (a restriction), per output row r:
//   x0 != nullptr: the output level's first Jacobi sweep from a zero guess,
//     JAC_REF: x0_r = r < nlim ? y_r / a_rr : 0 (a_rr = dv[dpos[r]], 0 without
//     a diagonal), JAC_L1: x0_r = y_r / l1_r -- k_jacobi0's arithmetic;
//   bd != nullptr: bd[dinv[r]] = y_r for rows with dinv[r] >= 0 (the dense
//     rows of the coarse direct solve, gathered to the dense numbering).
template <typename VT, typename XT>
struct Epi {
    XT* x0 = nullptr;
    const int* dpos = nullptr;
    const VT* dv = nullptr;
    const VT* l1 = nullptr;
    int nlim = 0;
    int kind = 0;
    double* bd = nullptr;
    const int* dinv = nullptr;
};

// lanes per row for a matrix with the given mean row length
inline int lanes_per_row(double mean_len) {
    if (mean_len <= 24) return 1;
    if (mean_len <= 48) return 4;
    if (mean_len <= 96) return 8;
    if (mean_len <= 192) return 16;
    return 32;
}

// returns the number of block partials written to part (if part != nullptr)
template <typename VT, typename XT>
int spmv(hipStream_t st, const DMat<VT>& A, const XT* x, const XT* b, XT* y, int mode, double* part, const XT* w,
         const Epi<VT, XT>* epi = nullptr);
// MODE_OUTER y = A x in FP64 with events recorded at the kernel's own start
// and end (the outer CG's timed launches)
int spmv_outer_timed(hipStream_t st, const DMat<double>& A, const double* x, double* y, double* part,
                     const double* w, hipEvent_t e0, hipEvent_t e1);
// block partials of a.b: at most DOT_PART_MAX of them (the grid's cap)
constexpr int DOT_PART_MAX = 1024;
template <typename XT>
int dot_part(hipStream_t st, int n, const XT* a, const XT* b, double* part);
// scal[slot] = sum(part); op 0 plain, 1 sqrt, 2 shift (scal[slot+1] = scal[slot]) then plain
void finalize(hipStream_t st, int np, const double* part, double* scal, int slot, int op);
// pcg's residual norm and stop test: unless scal[S_STOP] is already set,
// scal[S_R] = sqrt(sum(part)), scal[S_AMG0] = sum(part), scal[S_IT] += 1 and
// scal[S_STOP] = 1 when
// !((relres ? r : r / scal[S_R0]) > tol), r is NaN or S_IT reached max_iter
void finalize_stop(hipStream_t st, int np, const double* part, double* scal, double tol, int relres, int max_iter);
// Zero-guess preconditioner entry: block partials of F.F (k_dot_part's grid
// and order; returns their number); X (if given) = the initial guess, 0 and
// F_i / isa_i on the isolated rows (bit i of isb set; isb, isa: nullptr
// without isolated rows); xf (if given) = (XT) X; ff (if given) = (XT) F.
template <typename XT>
int precon_in(hipStream_t st, int n, const double* F, double* part, double* X, XT* xf, XT* ff,
              const uint32_t* isb, const double* isa);
// Zero-guess preconditioner exit: Z = the initial guess of precon_in when
// scal[slot] == 0, else the V-cycle's result (xs, or Z itself when xs is
// nullptr); w != nullptr: block partials of w.Z (k_dot_part's grid and order),
// returns their number
template <typename XT>
int precon_out(hipStream_t st, int n, const XT* xs, double* Z, const double* F, const uint32_t* isb,
               const double* isa, const double* scal, int slot, const double* w, double* part);
template <typename T>
void fill(hipStream_t st, int n, T* x, T val);
template <typename TI, typename TO>
void convert(hipStream_t st, int n, const TI* a, TO* b);
template <typename T>
void copy(hipStream_t st, int n, const T* a, T* b);
void signs(hipStream_t st, int n, const double* s, const double* x, double* y);
template <typename T>
void fix_isolated(hipStream_t st, int m, const int* rows, const double* a, const T* f, T* x);
void scale_by(hipStream_t st, int n, double* x, const double* scal, int slot, int inverse);
// batched Krylov over a block-diagonal system: chunks of rows that never
// straddle two systems (crow: nchunk+1 row starts, cinst: system of a chunk,
// ichunk: ninst+1 chunk starts), per-system scalars in device arrays
void seg_dot(hipStream_t st, int nchunk, const int* crow, const double* a, const double* b, double* part);
void seg_sum(hipStream_t st, int ninst, const int* ichunk, const double* part, double* out, double* old);
void bpcg_p(hipStream_t st, int nchunk, const int* crow, const int* cinst, double* P, const double* Z, const double* s,
            const double* sold, const double* active);
void bpcg_xr(hipStream_t st, int nchunk, const int* crow, const int* cinst, double* X, double* R, const double* P,
             const double* Q, const double* s, const double* pq, const double* active);
void bpcg_check(hipStream_t st, int ninst, const double* rr, const double* hb, double tol, int max_iter,
                double* active, double* iters, int count);
// reflective Newton box bookkeeping over the bounded index set (NaN bound: unbounded)
void rn_dj(hipStream_t st, int nb, const int* idx, const double* lo, const double* up, const double* G,
           const double* X, double* D, double* J);
void rn_reflect(hipStream_t st, int nb, const int* idx, const double* lo, const double* up, double* X);
void rn_snap(hipStream_t st, int nb, const int* idx, const double* lo, const double* up, double* X, double eps);
// M = D H D + diag(|g| J) over F = H's pattern plus the diagonal (fh: H has
// the position, fhv: its value), the values of the positions present in the
// cached pattern of M written to out[mpos]; *flag set when the presence of
// any position differs from the cache (flag is not cleared here; mpos ==
// nullptr: no cache); outF (optional): every F position's value, 0.0 where
// M has no entry
void rn_scaled_hessian(hipStream_t st, int n, const int* frp, const int* fci, const double* fhv, const uint8_t* fh,
                       const double* D, const double* G, const double* J, const int* mpos, double* out,
                       double* outF, int* flag);
// block partial counts of entries within 1e-30 of a bound; returns #partials
int rn_touch(hipStream_t st, int nb, const int* idx, const double* lo, const double* up, const double* X,
             double* part);
// dL/dtheta_e = -scale_e lam_e^T K_e(E=1) u_e, P1 plane strain (Poisson ratio nu)
void elast_grad(hipStream_t st, int ne, const int* tri, const double* xy, int nn, const double* u, const double* lam,
                double nu, const double* scale, int blocked, double* grad);
void mul(hipStream_t st, int n, double* x, const double* y);  // x_i *= y_i
void scale(hipStream_t st, int n, double* x, double a);       // x_i *= a
// when scal[slot] == 0: x = 0, then x_i = f_i / a_i on the isolated rows
// (the guess AMG_solve returns untouched when its first residual is zero)
void restore_if_zero(hipStream_t st, int n, double* x, const double* scal, int slot, int m, const int* rows,
                     const double* a, const double* f);
void pcg_p(hipStream_t st, int n, double* P, const double* Z, const double* scal, int first);
// (skips every update once scal[S_STOP] is set).  ff, xf: also the next
// zero-guess preconditioner entry on the new R, see k_pcg_xr
template <typename XT = double>
int pcg_xr(hipStream_t st, int n, double* X, double* R, const double* P, const double* Q, const double* scal,
           double* part, XT* ff = nullptr, XT* xf = nullptr, const double* sg = nullptr, const uint32_t* isb = nullptr,
           const double* isa = nullptr);
void axpy_s(hipStream_t st, int n, double* x, const double* y, const double* scal, int slot_a, int slot_b,
            double sign);
void axpby(hipStream_t st, int n, double a, double* x, double b, const double* y);
template <typename VT, typename XT>
void mcsor_colour(hipStream_t st, const DMat<VT>& Ac, int blk0, int nblk, const int* perm, const XT* f, XT* x,
                  XT omega);
// the first colour's sweep from a zero iterate (rows [0, rows0) of Ac), the
// other rows' entries of x zeroed -- the fill and the first colour launch in
// one elementwise pass
template <typename VT, typename XT>
void mcsor_first(hipStream_t st, const DMat<VT>& Ac, int rows0, const int* perm, const XT* f, XT* x, XT omega);
template <typename VT, typename XT>
void sor_lex(hipStream_t st, const DMat<VT>& A, int nlv, const int* lv_off, const int* lv_rows, const XT* f, XT* x,
             XT* xo, XT omega, int sweeps, int backward);
template <typename VT, typename XT>
void jacobi(hipStream_t st, const DMat<VT>& A, int kind, const XT* f, const XT* xo, XT* x, int nlim);
// the first sweep from x = 0, elementwise (no matrix row is read)
template <typename VT, typename XT>
void jacobi0(hipStream_t st, const DMat<VT>& A, int kind, const XT* f, XT* x, int nlim);
// in-place inverse of the N x N row-major matrix M (no pivoting), blocked
// by 64 pivots with an FP64 MFMA update; work: gauss_jordan_work(N) doubles
size_t gauss_jordan_work(int N);
void gauss_jordan_inverse(hipStream_t st, int N, double* M, double* work);
// block-diagonal dense inverse of a batch (block of dense row g: rblk[g];
// block b: rows [off[b], off[b+1]), matrix at M + moff[b]); piv: one per block
void batched_gauss_jordan_inverse(hipStream_t st, int nd, int max_block, const int* rblk, const int* off,
                                  const size_t* moff, double* M, double* piv);
// (active != nullptr: the dense rows of inactive systems -- block index =
// system index -- are skipped)
template <typename XT>
void bcoarse_solve(hipStream_t st, int nd, const int* rblk, const int* off, const size_t* moff, const int* didx,
                   const double* inv, int nt, const int* tidx, const double* tdiag, const XT* b, XT* x,
                   const double* active = nullptr);
// level-scheduled ILU(0) triangular solve (upper=0: L, diagonal last; 1: U, diagonal first)
void trsv(hipStream_t st, const int* rp, const int* ci, const double* v, int nlv, const int* lv_off,
          const int* lv_rows, const double* z, double* w, int upper);
// ILU(0) preconditioner w = U^-1 L^-1 z in one workgroup, both factors packed
// by level: phase order t (level by level), per row t its right-hand side
// in[t] and one record rec[t] (TriRec<maxp>, 48, 64 or 112 bytes) holding the
// diagonal, the first maxp off-diagonal values and their sources in the
// factor's row order, the output index and the off-diagonal count; entries
// past maxp come from ovval/ovsrc[ovp[t] ...].  A source >= 0 is a slot of
// the LDS ring of the last TRI_RING results (slot = t & (TRI_RING-1); slot
// TRI_RING holds 0 and pads the unused value slots, whose values are 0);
// -(g+1) reads out[g] in global memory (full[L]: level L reads so, and a
// full barrier precedes it; the flag is bit 31 of off[L]).  Row t's result
// goes to out[orow].
constexpr int TRI_NT = 1024;
constexpr int TRI_RING = 8192;
template <int MAXP>
struct alignas(16) TriRec {
    double diag;
    double val[MAXP];
    int src[MAXP];
    int orow;
    int cnt;
};
// rows of at most 3 off-diagonal entries, all sources in the ring (the
// 7-point stencils of the 2-D P1 factors): 48 bytes, 16-bit sources
template <>
struct alignas(16) TriRec<3> {
    double diag;
    double val[3];
    int orow;
    short src[3];
    short cnt;
    int pad;
};
static_assert(sizeof(TriRec<3>) == 48 && sizeof(TriRec<4>) == 64 && sizeof(TriRec<8>) == 112, "record layout");
struct TriPack {
    int n = 0, nlv = 0, maxp = 0;
    const int* off = nullptr;     // nlv + 1 level offsets, bit 31: full[L]
    const void* rec = nullptr;    // n records TriRec<maxp>
    const int* ovp = nullptr;     // n: first overflow entry of rows with cnt > maxp
    const double* ovval = nullptr;
    const int* ovsrc = nullptr;
};
// One exact lexicographic SOR sweep (k_sor_lex's arithmetic) over a matrix
// packed by the sweep's level schedule: p holds, per phase position, the
// diagonal and the entries whose values are the sweep's new ones (earlier
// rows in sweep order), with their sources as in TriPack and orow = the
// natural row; the entries whose values are the old ones (x before the
// sweep) are listed per position: ocnt of them, the first p.maxp in
// oval / ocol (slot-major), the rest from oovv / oovc[ooff[t] ...]; they
// enter as products computed before the sweep (prod slot-major, then prodov
// like oovv).  old_first: they come first in the rows (the backward sweep),
// else after the new ones.  New entries past p.maxp use p's overflow arrays.
struct SorPack {
    TriPack p;
    const int* row = nullptr;
    const int* ocnt = nullptr;
    const double* oval = nullptr;
    const int* ocol = nullptr;
    const int* ooff = nullptr;
    const double* oovv = nullptr;
    const int* oovc = nullptr;
    int old_first = 0;
    // scratch: n, n, maxp * n (zero when old_first), as many as oovv
    double *in = nullptr, *a = nullptr, *prod = nullptr, *prodov = nullptr;
};
void sor_packed(hipStream_t st, const SorPack& s, const double* f, double* x, double omega, int sweeps);
// forward phase fw: in = zp (z in L's phase order, gathered by perm), out = y
// (bw's phase order); backward phase bw: in = y, out = w (natural order)
void tri_solve_pair(hipStream_t st, const TriPack& fw, const TriPack& bw, const int* perm, const double* z,
                    double* zp, double* y, double* w);
// the packed values from the factor's CSR values: rec[t].val[k] =
// v[vmap[k*n + t]] (vmap < 0: 0), rec[t].diag = v[dmap[t]], ovval[q] = v[ovmap[q]]
void tri_pack_values(hipStream_t st, int n, int maxp, const int* vmap, const int* dmap, int nov, const int* ovmap,
                     const double* v, void* rec, double* ovval);
// The same factorisation from a symbolic plan (rows in level order t):
// m1[t] = {first A entry a0, A entries na, first L entry l0, U diagonal u0},
// m2[t] = {lower entries nl, A has the diagonal, first pair}, and per A entry
// p a bit mask mask[p] of the row's L slots q whose U(k_q, j) exists, their U
// entry indices listed in pu (slot order, entry after entry).  At most
// ILU_MAXA entries a row and ILU_MAXL lower ones.
constexpr int ILU_MAXA = 16, ILU_MAXL = 8;
// the planned factorisation one level per launch, each on a full grid (for
// schedules of few, wide levels; off: the host copy of the level offsets)
void ilu0_planned_levels(hipStream_t st, const std::vector<int>& off, const int4* m1, const int4* m2,
                         const int* mask, const int* pu, const double* vA, double* vL, double* vU, int* fail);
// a triangular sweep one level per launch (rows: the level-ordered rows of
// the factor, off: host level offsets); the arithmetic of PCG_preconl
// (upper = 0, diagonal last) or PCG_preconr (upper = 1, diagonal first)
void tri_levels_wide(hipStream_t st, const int* rp, const int* ci, const double* v, const std::vector<int>& off,
                     const int* rows, const double* in, double* out, int upper);
// out[r] = in[perm[r]] (inverse 0), out[perm[r]] = in[r] (inverse 1)
void permute(hipStream_t st, int n, const int* perm, const double* in, double* out, int inverse);
// out[k] = in[idx[k]], k < n (a gather of n entries)
void permute_idx(hipStream_t st, int n, const int* idx, const double* in, double* out);
void ilu0_planned(hipStream_t st, int nlv, const int* lv_off, const int4* m1, const int4* m2, const int* mask,
                  const int* pu, const double* vA, double* vL, double* vU, int* fail);
// level-scheduled ILU(0) numeric factorisation into fixed L/U patterns (see
// k_ilu0); A's rows sorted by column; *fail = 1 on a zero pivot
void ilu0_numeric(hipStream_t st, int nlv, const int* lv_off, const int* lv_rows, const int* rpA, const int* ciA,
                  const double* vA, const int* rpL, const int* ciL, double* vL, const int* rpU, const int* ciU,
                  double* vU, int* fail);
// out[l] = V_l . r, l < m (V column vectors with leading dimension ld); returns #partials per dot
int multidot(hipStream_t st, int n, int m, const double* V, size_t ld, const double* r, double* part, double* out);
// y = base + sign * sum_l V_l c_l
void combine(hipStream_t st, int n, int m, const double* V, size_t ld, const double* c, const double* base,
             double sign, double* y);
// bd: nd doubles of scratch for the gathered right-hand side; gathered:
// bd already holds it (a MODE_SETX restriction wrote it)
template <typename XT>
void coarse_solve(hipStream_t st, int nd, const int* didx, const double* inv, int nt, const int* tidx,
                  const double* tdiag, const XT* b, XT* x, double* bd, bool gathered = false);
// the same with the symmetric inverse's upper triangle as tiles of CT x CT
// values of type TT (FP32 or FP64; tile t = (ti[t], tj[t]), ti[t] <= tj[t],
// nct tiles a side, zero padded), FP64 products and sums; part: nct * nct *
// CT doubles of scratch, bd: nct * CT doubles, zero beyond nd
constexpr int CT = 64;
template <typename XT, typename TT>
void coarse_solve_sym(hipStream_t st, int nd, int nct, int ntiles, const int* ti, const int* tj, const TT* tiles,
                      double* part, const int* didx, int nt, const int* tidx, const double* tdiag, const XT* b, XT* x,
                      double* bd, bool gathered);
// tiles = the upper triangle of inv (nd x nd, FP64) as TT, symmetrised
template <typename TT>
void inverse_to_tiles(hipStream_t st, int nd, int ntiles, const int* ti, const int* tj, const double* inv,
                      TT* tiles);

// Galerkin products from a plan (rap.cpp): out[e] = the products va[ia[t]] *
// vb[ib[t]], t in [ptr[e], ptr[e+1]), summed in that order, from +0.0 when
// seed[e], else from the first product
void plan_sum(hipStream_t st, int ne, const int64_t* ptr, const int* ia, const int* ib, const uint8_t* seed,
              const double* va, const double* vb, double* out);
// the coarse entries into the level's CSR values: dst[pos[e]] = val[e];
// *flag = 1 when an entry's zero / nonzero state differs from the pattern
// (pos < 0: not in it; lone: a zero there is the pattern's single (i, 0.0))
void plan_scatter(hipStream_t st, int ne, const double* val, const int* pos, const uint8_t* lone, double* dst,
                  int* flag);
// symmetric Dirichlet treatment of a CSR system keeping its pattern (dir[i]:
// DOF i constrained to u[i]): constrained rows become unit rows with
// b[i] = u[i]; every other row r gets b[r] -= sum over its constrained
// columns c (column order, from +0.0) of a_rc u_c, and those a_rc = 0.  b
// may be null (matrix only)
void dirichlet_rows(hipStream_t st, int n, const int* rp, const int* ci, const uint8_t* dir, const double* u,
                    double* val, double* b);
// val = -val on the m listed rows
void negate_rows(hipStream_t st, int m, const int* rows, const int* rp, double* v);
// out[r] = sum_p |(VT) v[p]| over row r, in VT arithmetic and column order
template <typename VT>
void row_l1(hipStream_t st, int n, const int* rp, const double* v, VT* out);
// dst[i] = idx[i] >= 0 ? src[idx[i]] : 0
template <typename T>
void gather(hipStream_t st, int n, const int* idx, const T* src, T* dst);
// dst[idx[i]] = src[i] where idx[i] >= 0
void scatter(hipStream_t st, int n, const int* idx, const double* src, double* dst);

}  // namespace dev
}  // namespace lina
