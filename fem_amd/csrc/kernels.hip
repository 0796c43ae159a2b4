// Device kernels of liblina_hip (gfx950 / CDNA4, wave64).
//
// Row kernels (SpMV, multicolour SOR, Jacobi) share one structure:
//
//   phase 1  all 256 lanes of a workgroup stream the value and column arrays
//            of a block of consecutive rows (<= 2048 entries, coalesced, each
//            matrix entry read once per sweep) and store the products
//            a_ij * x_j in LDS.  A block's bounds come as one (first row,
//            first entry) pair, and each lane group's row data is loaded
//            together with the streamed entries, so a block waits for two
//            memory round trips, not five;
//   phase 2  rows are reduced from LDS by groups of LPR lanes.  LPR == 1
//            adds a row's products one after another in column order, the
//            order of the reference's `sum += A->Values[i][j]*x[...]` loops
//            (LinA/linear_algebra.c:134-184), so the result is bit-identical;
//            LPR > 1 (the dense Galerkin coarse levels, 50-200 entries a row)
//            adds lane partials with a fixed xor-shuffle tree.
//
// A row longer than a block gets a block of its own and is reduced by the
// whole workgroup.  The library is compiled with -ffp-contract=off, so no
// multiply-add is fused.  Reductions (dots, norms) are deterministic: fixed
// per-block trees, block partials stored by logical block index, a
// single-block finalisation.

#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <cstring>

#include <algorithm>
#include <cstdint>
#include <cstdlib>

#include "kernels.hpp"

#include <stdexcept>
#include <string>

namespace lina {
namespace dev {

constexpr int TPB = 256;  // threads per block of the row kernels
constexpr int WAVE = 64;

// Workgroups are dispatched round-robin over the 8 XCDs (each with its own
// L2).  Remap so that XCD x processes one contiguous range of logical blocks:
// neighbouring rows (and the x entries they gather) then share an L2.
__device__ __forceinline__ int xcd_block(int bid, int nblk) {
    constexpr int X = 8;
    if (nblk < 2 * X) return bid;
    int q = nblk / X, r = nblk % X;
    int x = bid % X, k = bid / X;
    return x * q + min(x, r) + k;
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
    for (int o = WAVE / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, WAVE);
    return v;
}

template <int LPR, typename T>
__device__ __forceinline__ T group_sum(T v) {
    for (int o = LPR / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, LPR);
    return v;
}

// deterministic block sum (fixed shuffle tree, then waves summed in order)
__device__ __forceinline__ double block_sum(double v) {
    __shared__ double ws[TPB / WAVE];
    v = wave_sum(v);
    int lane = threadIdx.x % WAVE, w = threadIdx.x / WAVE;
    __syncthreads();
    if (lane == 0) ws[w] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int i = 0; i < TPB / WAVE; i++) s += ws[i];
    return s;  // valid in thread 0
}

#ifndef LINA_NT_LOADS
#define LINA_NT_LOADS 1
#endif

// Loads of streamed arrays.  NT: nontemporal (the FP64 matrices, whose
// entries are read once per sweep and would otherwise displace the gathered
// vector from L2); FP32 matrices measured faster with ordinary loads.
template <bool NT, typename T>
__device__ __forceinline__ T ld_stream(const T* p) {
    if constexpr (NT && LINA_NT_LOADS) return __builtin_nontemporal_load(p);
    else return *p;
}

// phase 1, split in two so that a kernel can issue its own per-row loads
// between them: load() issues every column and value load of the lane's
// entries k = k0 + lane + u*TPB (U = block_nnz/TPB of them), finish() issues
// the gathers x[ci[k]] and stores the products v[k] * x[ci[k]] in LDS at
// k - k0.  Up to 3U loads of a lane are in flight at once.
#ifndef LINA_NT_FP32
#define LINA_NT_FP32 0
#endif
template <typename VT, typename XT>
struct Stage {
    static constexpr int U = block_nnz<VT>() / TPB;
    static constexpr bool NT = sizeof(VT) == 8 || LINA_NT_FP32;
    int c[U];
    VT a[U];
    // entries past k1 load entry k1-1 again (in bounds, same cache line) so
    // that every load is issued unconditionally; k1 > k0.  cb.x >= 0: the
    // block's 16-bit column codes (see DMat::cs) with bases cb.
    __device__ __forceinline__ void load(int k0, int k1, const int* __restrict__ ci,
                                         const uint16_t* __restrict__ cs, int2 cb, const VT* __restrict__ v) {
        const int base = k0 + threadIdx.x;
        if (cs && cb.x >= 0) {
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int k = min(base + u * TPB, k1 - 1);
                const int o = (int)ld_stream<NT>(cs + k);
                c[u] = ((o & 0x8000) ? cb.y : cb.x) + (o & 0x7fff);
                a[u] = ld_stream<NT>(v + k);
            }
        } else {
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int k = min(base + u * TPB, k1 - 1);
                c[u] = ld_stream<NT>(ci + k);
                a[u] = ld_stream<NT>(v + k);
            }
        }
    }
    __device__ __forceinline__ void finish(int k0, int k1, const XT* __restrict__ x, XT* prod) {
        const int base = k0 + threadIdx.x;
        XT xv[U];
#pragma unroll
        for (int u = 0; u < U; u++) xv[u] = x[c[u]];
#pragma unroll
        for (int u = 0; u < U; u++)
            if (base + u * TPB < k1) prod[base + u * TPB - k0] = (XT)a[u] * xv[u];
    }
};

// ------------------------------------------------------------------ SpMV
//
// MODE_SET : y = A x                        (sparse_mult / PCG_mult)
// MODE_RES : y = b + (-(A x))               (linear_map(R,-1,A,X) on R=b)
// MODE_ADD : y = y + (A x)                  (linear_map(X, 1,P,c))
// MODE_ADDB: y = b + (A x)                  (the same sum written beside b)
// MODE_SETX: y = A x, then the Epi epilogue on row r
// part != nullptr: block partial of sum_r y_r * w_r (w == nullptr: y_r^2)

// the vector entry a row's result is combined with (b_r or y_r), if any
template <int MODE, typename XT>
__device__ __forceinline__ XT spmv_base(int r, const XT* b, const XT* y) {
    if (MODE == MODE_RES || MODE == MODE_ADDB) return b[r];
    if (MODE == MODE_ADD) return y[r];
    return (XT)0;
}

template <typename VT, typename XT, int MODE>
__device__ __forceinline__ XT spmv_out(XT s, XT base, int r, XT* y, const Epi<VT, XT>& e) {
    XT out;
    if (MODE == MODE_SET || MODE == MODE_OUTER || MODE == MODE_SETX) out = s;
    else if (MODE == MODE_RES) out = base + (-s);
    else out = base + s;
    y[r] = out;
    if (MODE == MODE_SETX) {
        if (e.x0) {
            if (e.kind == JAC_REF) {
                if (r >= e.nlim) {
                    e.x0[r] = (XT)0;
                } else {
                    const int dp = e.dpos[r];
                    e.x0[r] = out / (dp >= 0 ? (XT)e.dv[dp] : (XT)0);
                }
            } else {
                e.x0[r] = out / (XT)e.l1[r];
            }
        }
        if (e.bd) {
            const int g = e.dinv[r];
            if (g >= 0) e.bd[g] = (double)out;
        }
    }
    return out;
}

// Logical row block of iteration k of workgroup b.  A launch with one
// workgroup per block (G >= nblk) takes xcd_block's remap; a persistent grid
// (G < nblk, G a multiple of 8) gives XCD x the contiguous block range
// [x*chunk, (x+1)*chunk) and walks it with the XCD's G/8 workgroups, so that
// workgroups stay resident instead of being dispatched once per block.
__device__ __forceinline__ int grid_block(int b, int G, int k, int nblk) {
    if (G >= nblk) return k == 0 ? xcd_block(b, nblk) : -1;
    constexpr int X = 8;
    const int per = G / X, x = b % X, j = b / X;
    const int chunk = (nblk + X - 1) / X;
    const int local = j + k * per;
    if (local >= chunk) return -1;
    const int idx = x * chunk + local;
    return idx < nblk ? idx : -1;
}

// a batch's row block is live while one of its systems is active (always,
// until the flags are attached: the autotuner runs during upload)
__device__ __forceinline__ bool block_live(const int2* __restrict__ bspec, const double* __restrict__ active, int bb) {
    if (!bspec || !active) return true;
    const int2 sp = bspec[bb];
    for (int k = sp.x; k <= sp.y; k++)
        if (active[k] != 0.0) return true;
    return false;
}

// One row block bb.  The first row of each lane group's reduction (all rows
// of a block when it has at most TPB/LPR rows) has its row bounds and base
// entry loaded together with the streamed entries, off the critical path.
template <typename VT, typename XT, int MODE, int LPR>
__device__ __forceinline__ void spmv_block(int bb, const int2* __restrict__ bk, const int* __restrict__ rp,
                                           const int* __restrict__ ci, const uint16_t* __restrict__ cs,
                                           const int2* __restrict__ cbase, const VT* __restrict__ v,
                                           const XT* __restrict__ x, const XT* __restrict__ b, XT* __restrict__ y,
                                           double* __restrict__ part, const XT* __restrict__ w,
                                           const Epi<VT, XT>& epi, XT* prod) {
    constexpr int BN = block_nnz<VT>();
    const int2 s0 = bk[bb], s1 = bk[bb + 1];
    const int r0 = s0.x, r1 = s1.x, k0 = s0.y, k1 = s1.y;
    double mine = 0.0;
    if (k1 - k0 > BN) {
        // one long row: strided partial sums + block tree
        XT s = 0;
        for (int k = k0 + threadIdx.x; k < k1; k += TPB) s += (XT)v[k] * x[ci[k]];
        double t = block_sum((double)s);
        if (threadIdx.x == 0) {
            XT out = spmv_out<VT, XT, MODE>((XT)t, spmv_base<MODE>(r0, b, y), r0, y, epi);
            if (part) mine = (double)out * (double)(w ? w[r0] : out);
        }
    } else {
        constexpr int G = TPB / LPR;
        const int g = threadIdx.x / LPR, lane = threadIdx.x % LPR;
        const int rf = r0 + g;
        Stage<VT, XT> st;
        if (k1 > k0) st.load(k0, k1, ci, cs, cs ? cbase[bb] : make_int2(-1, -1), v);
        int pb = 0, pe = 0;
        XT base0 = (XT)0;
        if (rf < r1) {
            pb = rp[rf];
            pe = rp[rf + 1];
            base0 = spmv_base<MODE>(rf, b, y);
        }
        if (k1 > k0) st.finish(k0, k1, x, prod);
        __syncthreads();
        for (int r = rf; r < r1; r += G) {
            const bool first = r == rf;
            XT s = 0;
            const int e = (first ? pe : rp[r + 1]) - k0;
            for (int k = (first ? pb : rp[r]) - k0 + lane; k < e; k += LPR) s += prod[k];
            if (LPR > 1) s = group_sum<LPR>(s);
            if (lane == 0) {
                XT out = spmv_out<VT, XT, MODE>(s, first ? base0 : spmv_base<MODE>(r, b, y), r, y, epi);
                if (part) mine += (double)out * (double)(w ? w[r] : out);
            }
        }
    }
    if (part) {
        double t = block_sum(mine);
        if (threadIdx.x == 0) part[bb] = t;
    }
}

template <typename VT, typename XT, int MODE, int LPR>
__global__ __launch_bounds__(TPB) void k_spmv(int nblk, const int2* __restrict__ bk, const int* __restrict__ rp,
                                              const int* __restrict__ ci, const uint16_t* __restrict__ cs,
                                              const int2* __restrict__ cbase, const VT* __restrict__ v,
                                              const XT* __restrict__ x, const XT* __restrict__ b,
                                              XT* __restrict__ y, double* __restrict__ part,
                                              const XT* __restrict__ w, const Epi<VT, XT> epi,
                                              const int2* __restrict__ bspec, const double* __restrict__ active) {
    __shared__ XT prod[block_nnz<VT>()];
    for (int it = 0;; it++) {
        const int bb = grid_block(blockIdx.x, gridDim.x, it, nblk);
        if (bb < 0) break;
        if (!block_live(bspec, active, bb)) continue;
        spmv_block<VT, XT, MODE, LPR>(bb, bk, rp, ci, cs, cbase, v, x, b, y, part, w, epi, prod);
        __syncthreads();  // prod is restaged by the next block
    }
}

// ------------------------------------------------------------ reductions

template <typename XT>
__global__ __launch_bounds__(TPB) void k_dot_part(int n, const XT* __restrict__ a, const XT* __restrict__ b,
                                                  double* __restrict__ part) {
    double s = 0.0;
    for (int i = blockIdx.x * TPB + threadIdx.x; i < n; i += gridDim.x * TPB) s += (double)a[i] * (double)b[i];
    double t = block_sum(s);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// Sum of np block partials by one workgroup of FTPB threads: four
// independent accumulators per thread (so that four loads are in flight),
// then the wave trees, then the waves in order.  Deterministic.
constexpr int FTPB = 1024;
__device__ __forceinline__ double partial_sum(int np, const double* __restrict__ part) {
    __shared__ double ws[FTPB / WAVE];
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    int i = threadIdx.x;
    for (; i + 3 * FTPB < np; i += 4 * FTPB) {
        s0 += part[i];
        s1 += part[i + FTPB];
        s2 += part[i + 2 * FTPB];
        s3 += part[i + 3 * FTPB];
    }
    for (; i < np; i += FTPB) s0 += part[i];
    double v = wave_sum((s0 + s1) + (s2 + s3));
    if (threadIdx.x % WAVE == 0) ws[threadIdx.x / WAVE] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < FTPB / WAVE; w++) t += ws[w];
    return t;  // valid in thread 0
}

// out[slot] = sum(part); op: 0 plain, 1 sqrt, 2 shift (out[slot+1]=out[slot]) then plain
__global__ __launch_bounds__(FTPB) void k_finalize(int np, const double* __restrict__ part, double* __restrict__ scal,
                                                   int slot, int op) {
    double t = partial_sum(np, part);
    if (threadIdx.x == 0) {
        if (op == 2) scal[slot + 1] = scal[slot];
        scal[slot] = (op == 1) ? sqrt(t) : t;
    }
}

// pcg's |R| and stop test (the loop condition of PCGS.c:224 evaluated on the
// device): nothing happens once scal[S_STOP] is set
__global__ __launch_bounds__(FTPB) void k_finalize_stop(int np, const double* __restrict__ part,
                                                        double* __restrict__ scal, double tol, int relres,
                                                        int max_iter) {
    if (scal[S_STOP] != 0.0) return;
    double t = partial_sum(np, part);
    if (threadIdx.x == 0) {
        const double r = sqrt(t);
        scal[S_R] = r;
        scal[S_AMG0] = t;  // the next zero-guess preconditioner's r0 test when pcg_xr prepared its entry
        const double it = scal[S_IT] + 1.0;
        scal[S_IT] = it;
        const double crit = relres ? r : r / scal[S_R0];
        if (!(crit > tol) || r != r || it >= (double)max_iter) scal[S_STOP] = 1.0;
    }
}

// row i is isolated (bit i of the mask)
__device__ __forceinline__ bool isolated(const uint32_t* __restrict__ isb, int i) {
    return isb && ((isb[i >> 5] >> (i & 31)) & 1u);
}

// zero-guess preconditioner entry (see precon_in); the F.F partials follow
// k_dot_part's loop and tree exactly
template <typename XT>
__global__ __launch_bounds__(TPB) void k_precon_in(int n, const double* __restrict__ F, double* __restrict__ part,
                                                   double* __restrict__ X, XT* __restrict__ xf, XT* __restrict__ ff,
                                                   const uint32_t* __restrict__ isb, const double* __restrict__ isa) {
    double s = 0.0;
    for (int i = blockIdx.x * TPB + threadIdx.x; i < n; i += gridDim.x * TPB) {
        const double f = F[i];
        s += f * f;
        if (X || xf) {
            double x = 0.0;
            if (isolated(isb, i)) x = f / isa[i];
            if (X) X[i] = x;
            if (xf) xf[i] = (XT)x;
        }
        if (ff) ff[i] = (XT)f;
    }
    double t = block_sum(s);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// zero-guess preconditioner exit (see precon_out)
template <typename XT>
__global__ __launch_bounds__(TPB) void k_precon_out(int n, const XT* __restrict__ xs, double* __restrict__ Z,
                                                    const double* __restrict__ F, const uint32_t* __restrict__ isb,
                                                    const double* __restrict__ isa,
                                                    const double* __restrict__ scal, int slot,
                                                    const double* __restrict__ w, double* __restrict__ part) {
    const bool restore = scal[slot] == 0.0;
    double s = 0.0;
    for (int i = blockIdx.x * TPB + threadIdx.x; i < n; i += gridDim.x * TPB) {
        double z;
        if (restore) {
            z = 0.0;
            if (isolated(isb, i)) z = F[i] / isa[i];
            Z[i] = z;
        } else if (xs) {
            z = (double)xs[i];
            Z[i] = z;
        } else {
            z = Z[i];
        }
        if (w) s += w[i] * z;
    }
    if (w) {
        double t = block_sum(s);
        if (threadIdx.x == 0) part[blockIdx.x] = t;
    }
}

// ---------------------------------------------------------- vector ops

template <typename T>
__global__ void k_fill(int n, T* __restrict__ x, T val) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) x[i] = val;
}

template <typename TI, typename TO>
__global__ void k_convert(int n, const TI* __restrict__ a, TO* __restrict__ b) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) b[i] = (TO)a[i];
}

template <typename T>
__global__ void k_copy(int n, const T* __restrict__ a, T* __restrict__ b) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) b[i] = a[i];
}

// y = s .* x (sparse_multiplication with the diagonal Signs matrix: 0 + s*x)
__global__ void k_signs(int n, const double* __restrict__ s, const double* __restrict__ x, double* __restrict__ y) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) y[i] = 0.0 + s[i] * x[i];
}

// AMG_fix_isolated (Multigrid.c:1301-1319): Sol[i] = F[i]/a
template <typename T>
__global__ void k_fix_isolated(int m, const int* __restrict__ rows, const double* __restrict__ a,
                               const T* __restrict__ f, T* __restrict__ x) {
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < m; k += gridDim.x * blockDim.x) {
        int i = rows[k];
        x[i] = (T)((double)f[i] / a[k]);
    }
}

__global__ void k_zero_if(int n, double* __restrict__ x, const double* __restrict__ scal, int slot) {
    if (scal[slot] != 0.0) return;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) x[i] = 0.0;
}

__global__ void k_fix_isolated_if(int m, const int* __restrict__ rows, const double* __restrict__ a,
                                  const double* __restrict__ f, double* __restrict__ x,
                                  const double* __restrict__ scal, int slot) {
    if (scal[slot] != 0.0) return;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < m; k += gridDim.x * blockDim.x) {
        int i = rows[k];
        x[i] = f[i] / a[k];
    }
}

// scale by 1/h (pcg: scalar_mult((double)1/h, ...)), h read from scal[slot]
// (h==0 -> 1); inverse==0 scales by h itself (the final unscaling)
__global__ void k_scale_by(int n, double* __restrict__ x, const double* __restrict__ scal, int slot, int inverse) {
    double h = scal[slot];
    if (h == 0) h = 1;
    double f = inverse ? (double)1 / h : h;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) x[i] *= f;
}

// pcg P update: i==0: P = Z; else P = (P*b) + Z with b = s/s_old
__global__ void k_pcg_p(int n, double* __restrict__ P, const double* __restrict__ Z, const double* __restrict__ scal,
                        int first) {
    if (first) {
        for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) P[i] = Z[i];
    } else {
        double b = scal[S_S] / scal[S_SOLD];
        for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
            double p = P[i] * b;
            P[i] = p + 1 * Z[i];
        }
    }
}

// pcg: a = s/(P.Q); X += a P; R += (-a) Q; partial |R|^2.  Also the next
// zero-guess preconditioner entry on the new R where asked (precon_in's
// outputs, with the signed right-hand side ft = 0 + sg_i r_i when sg is
// given): ff = (XT) ft when ff is given, xf = the initial guess (0; ft / isa_i
// on the rows marked in isb) when xf is given.  Sum of squares of ft = that
// of r.
template <typename XT>
__global__ __launch_bounds__(TPB) void k_pcg_xr(int n, double* __restrict__ X, double* __restrict__ R,
                                                const double* __restrict__ P, const double* __restrict__ Q,
                                                const double* __restrict__ scal, double* __restrict__ part,
                                                XT* __restrict__ ff, XT* __restrict__ xf,
                                                const double* __restrict__ sg, const uint32_t* __restrict__ isb,
                                                const double* __restrict__ isa) {
    if (scal[S_STOP] != 0.0) return;
    double a = scal[S_S] / scal[S_PQ];
    double na = -a;
    double s = 0.0;
    for (int i = blockIdx.x * TPB + threadIdx.x; i < n; i += gridDim.x * TPB) {
        X[i] += a * P[i];
        double r = R[i] + na * Q[i];
        R[i] = r;
        s += r * r;
        if (ff || xf) {
            const double ft = sg ? 0.0 + sg[i] * r : r;
            if (ff) ff[i] = (XT)ft;
            if (xf) {
                double x0 = 0.0;
                if (isolated(isb, i)) x0 = ft / isa[i];
                xf[i] = (XT)x0;
            }
        }
    }
    double t = block_sum(s);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// generic x += f*y with f = sign * scal[slot_a] / scal[slot_b] (slot_b < 0: 1)
__global__ void k_axpy_s(int n, double* __restrict__ x, const double* __restrict__ y, const double* __restrict__ scal,
                         int slot_a, int slot_b, double sign) {
    double f = scal[slot_a];
    if (slot_b >= 0) f = f / scal[slot_b];
    f = sign * f;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) x[i] += f * y[i];
}

// x = a*x + b*y with scalar constants (host-known)
__global__ void k_axpby(int n, double a, double* __restrict__ x, double b, const double* __restrict__ y) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) x[i] = a * x[i] + b * y[i];
}

// ------------------------------------------------------------- smoothers

// Multicolour SOR, one colour: the rows of the colour-permuted copy in the
// block range [blk0, blk0+nblk); row r of the copy is original row perm[r].
// x_i = (1-w) x_i + w*(f_i - sum_{j!=i} a_ij x_j)/a_ii with the arithmetic of
// AMG_SOR_smoother (Multigrid.c:1345-1365).  Rows of one colour are not
// coupled, so the products read x before any row of the colour is updated.
template <typename VT, typename XT, int LPR>
__global__ __launch_bounds__(TPB) void k_mcsor(int nblk, int blk0, const int2* __restrict__ bk,
                                               const int* __restrict__ rp, const int* __restrict__ ci,
                                               const uint16_t* __restrict__ cs, const int2* __restrict__ cbase,
                                               const VT* __restrict__ v, const int* __restrict__ dpos,
                                               const int* __restrict__ perm, const XT* __restrict__ f,
                                               XT* __restrict__ x, XT omega, const int2* __restrict__ bspec,
                                               const double* __restrict__ active) {
    constexpr int BN = block_nnz<VT>();
    __shared__ XT prod[BN];
    const int bb = blk0 + xcd_block(blockIdx.x, nblk);
    if (!block_live(bspec, active, bb)) return;
    const int2 s0 = bk[bb], s1 = bk[bb + 1];
    const int r0 = s0.x, r1 = s1.x, k0 = s0.y, k1 = s1.y;
    if (k1 - k0 > BN) {
        // one long row: the whole workgroup reduces it (block tree)
        const int i = perm[r0];
        const int dp = dpos[r0];
        double s = 0.0;
        for (int k = k0 + threadIdx.x; k < k1; k += TPB)
            if (k != dp) s += (double)((XT)v[k] * x[ci[k]]);
        s = block_sum(s);
        if (threadIdx.x == 0) {
            const XT d = dp >= 0 ? (XT)v[dp] : (XT)0;
            x[i] = ((XT)1 - omega) * x[i] + omega * (f[i] - (XT)s) / d;
        }
        return;
    }
    constexpr int G = TPB / LPR;
    const int g = threadIdx.x / LPR, lane = threadIdx.x % LPR;
    const bool leader = LPR == 1 || lane == 0;
    // the group's first row: permutation, bounds and diagonal position first,
    // then the streamed entries, then f_i, x_i and a_ii (x_i is written only
    // by this lane, after every gather of the block)
    const int rf = r0 + g;
    int pi = 0, pb = 0, pe = 0, pdp = -1;
    if (rf < r1) {
        pi = perm[rf];
        pb = rp[rf];
        pe = rp[rf + 1];
        pdp = dpos[rf];
    }
    Stage<VT, XT> st;
    if (k1 > k0) st.load(k0, k1, ci, cs, cs ? cbase[bb] : make_int2(-1, -1), v);
    XT pf = 0, px = 0, pd = 0;
    if (rf < r1 && leader) {
        pf = f[pi];
        px = x[pi];
        pd = pdp >= 0 ? (XT)v[pdp] : (XT)0;
    }
    if (k1 > k0) st.finish(k0, k1, x, prod);
    __syncthreads();
    for (int r = rf; r < r1; r += G) {
        const bool first = r == rf;
        const int i = first ? pi : perm[r];
        const int kb = first ? pb : rp[r];
        const int e = first ? pe : rp[r + 1];
        const int dp = first ? pdp : dpos[r];
        if (LPR == 1) {
            XT sum = first ? pf : f[i];
            for (int k = kb; k < e; k++)
                if (k != dp) sum -= prod[k - k0];
            const XT d = first ? pd : (dp >= 0 ? (XT)v[dp] : (XT)0);
            const XT xi = first ? px : x[i];
            x[i] = ((XT)1 - omega) * xi + omega * sum / d;
        } else {
            XT s = 0;
            for (int k = kb + lane; k < e; k += LPR)
                if (k != dp) s += prod[k - k0];
            s = group_sum<LPR>(s);
            if (lane == 0) {
                const XT d = first ? pd : (dp >= 0 ? (XT)v[dp] : (XT)0);
                const XT fi = first ? pf : f[i];
                const XT xi = first ? px : x[i];
                x[i] = ((XT)1 - omega) * xi + omega * (fi - s) / d;
            }
        }
    }
}

// Exact lexicographic SOR (the reference's AMG_SOR_smoother) by level
// scheduling inside one workgroup: rows of one dependency level are
// independent; j<i reads the new x, j>i the snapshot xo of the sweep start.
// backward != 0: the reversed sweep (rows n-1..0), j>i new, j<i snapshot.
template <typename VT, typename XT>
__global__ __launch_bounds__(1024) void k_sor_lex(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                                                  const VT* __restrict__ v, int nlv, const int* __restrict__ lv_off,
                                                  const int* __restrict__ lv_rows, const XT* __restrict__ f,
                                                  XT* __restrict__ x, XT* __restrict__ xo, XT omega, int sweeps,
                                                  int backward) {
    for (int s = 0; s < sweeps; s++) {
        for (int i = threadIdx.x; i < n; i += blockDim.x) xo[i] = x[i];
        __syncthreads();
        for (int L = 0; L < nlv; L++) {
            const int e = lv_off[L + 1];
            for (int t = lv_off[L] + threadIdx.x; t < e; t += blockDim.x) {
                const int i = lv_rows[t];
                XT sum = f[i];
                XT d = 0;
                for (int k = rp[i]; k < rp[i + 1]; k++) {
                    int j = ci[k];
                    if (j == i) d = (XT)v[k];
                    else if ((j < i) != (backward != 0)) sum -= (XT)v[k] * x[j];
                    else sum -= (XT)v[k] * xo[j];
                }
                x[i] = ((XT)1 - omega) * x[i] + omega * sum / d;
            }
            __syncthreads();
        }
    }
}

// Jacobi sweep x = J(xo) (x and xo distinct):
//   JAC_REF : x_i = (f_i - sum_{j!=i} a_ij xo_j)/a_ii for i < nlim, x_i = xo_i
//             beyond (AMG_parallel_Jacobi_smoother, Multigrid.c:1367-1400)
//   JAC_L1  : x_i = xo_i + (f_i - sum_j a_ij xo_j)/sum_j |a_ij|
template <typename VT, typename XT, int KIND, int LPR>
__global__ __launch_bounds__(TPB) void k_jacobi(int nblk, const int2* __restrict__ bk, const int* __restrict__ rp,
                                                const int* __restrict__ ci, const uint16_t* __restrict__ cs,
                                                const int2* __restrict__ cbase, const VT* __restrict__ v,
                                                const int* __restrict__ dpos, const VT* __restrict__ l1,
                                                const XT* __restrict__ f, const XT* __restrict__ xo,
                                                XT* __restrict__ x, int nlim, const int2* __restrict__ bspec,
                                                const double* __restrict__ active) {
    constexpr int BN = block_nnz<VT>();
    __shared__ XT prod[BN];
    const int bb = xcd_block(blockIdx.x, nblk);
    if (!block_live(bspec, active, bb)) return;
    const int2 s0 = bk[bb], s1 = bk[bb + 1];
    const int r0 = s0.x, r1 = s1.x, k0 = s0.y, k1 = s1.y;
    if (k1 - k0 > BN) {
        // one long row: the whole workgroup reduces it (block tree)
        const int r = r0;
        if (KIND == JAC_REF && r >= nlim) {
            if (threadIdx.x == 0) x[r] = xo[r];
            return;
        }
        const int dp = KIND == JAC_REF ? dpos[r] : -1;
        double s = 0.0;
        for (int k = k0 + threadIdx.x; k < k1; k += TPB)
            if (k != dp) s += (double)((XT)v[k] * xo[ci[k]]);
        s = block_sum(s);
        if (threadIdx.x == 0) {
            const XT rr = f[r] - (XT)s;
            if (KIND == JAC_REF) x[r] = rr / (dp >= 0 ? (XT)v[dp] : (XT)0);
            else x[r] = xo[r] + rr / l1[r];
        }
        return;
    }
    constexpr int G = TPB / LPR;
    const int g = threadIdx.x / LPR, lane = threadIdx.x % LPR;
    const bool leader = LPR == 1 || lane == 0;
    // the group's first row: bounds and (l1-Jacobi) f_r, xo_r, l1_r loaded
    // with the streamed entries
    const int rf = r0 + g;
    int pb = 0, pe = 0;
    XT pf = 0, pxo = 0;
    VT pl1 = 0;
    Stage<VT, XT> st;
    if (k1 > k0) st.load(k0, k1, ci, cs, cs ? cbase[bb] : make_int2(-1, -1), v);
    if (rf < r1) {
        pb = rp[rf];
        pe = rp[rf + 1];
        if (KIND == JAC_L1 && leader) {
            pf = f[rf];
            pxo = xo[rf];
            pl1 = l1[rf];
        }
    }
    if (k1 > k0) st.finish(k0, k1, xo, prod);
    __syncthreads();
    for (int r = rf; r < r1; r += G) {
        const bool first = r == rf;
        const int kb = first ? pb : rp[r];
        const int e = first ? pe : rp[r + 1];
        if (KIND == JAC_REF && r >= nlim) {
            if (lane == 0) x[r] = xo[r];  // rows past (n/dof)*dof are not swept
            continue;
        }
        const int dp = KIND == JAC_REF ? dpos[r] : -1;
        const XT fr = (KIND == JAC_L1 && first) ? pf : f[r];
        XT s;
        if (LPR == 1) {
            s = fr;
            for (int k = kb; k < e; k++)
                if (k != dp) s -= prod[k - k0];
        } else {
            s = 0;
            for (int k = kb + lane; k < e; k += LPR)
                if (k != dp) s += prod[k - k0];
            s = fr - group_sum<LPR>(s);
        }
        if (lane == 0) {
            if (KIND == JAC_REF) x[r] = s / (dp >= 0 ? (XT)v[dp] : (XT)0);
            else x[r] = (first ? pxo : xo[r]) + s / (first ? pl1 : l1[r]);
        }
    }
}

// First Jacobi sweep from x = 0 (the coarse levels of a V-cycle start from a
// zero guess): every product vanishes, so the sweep is elementwise and reads
// no matrix row, with the same result as k_jacobi on xo = 0:
//   JAC_REF : x_i = f_i / a_ii for i < nlim, x_i = 0 beyond
//   JAC_L1  : x_i = f_i / sum_j |a_ij|
template <typename VT, typename XT, int KIND>
__global__ void k_jacobi0(int n, const int* __restrict__ dpos, const VT* __restrict__ v, const VT* __restrict__ l1,
                          const XT* __restrict__ f, XT* __restrict__ x, int nlim) {
    for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x) {
        if (KIND == JAC_REF) {
            if (r >= nlim) {
                x[r] = (XT)0;
            } else {
                const int dp = dpos[r];
                x[r] = f[r] / (dp >= 0 ? (XT)v[dp] : (XT)0);
            }
        } else {
            x[r] = f[r] / (XT)l1[r];
        }
    }
}

// ------------------------------------------ ILU(0) triangular solves
//
// PCG_preconl / PCG_preconr (LinA/PCGS.c:61-100) by level scheduling inside
// one workgroup.  Forward: L rows hold the diagonal last,
//   w_i = (z_i - sum_{j<last} l_ij w_j) / l_ii;
// backward: U rows hold the (unit) diagonal first,
//   w_i = (y_i - sum_{j>first} u_ij w_j) / u_ii.
// Rows of one level depend only on rows of earlier levels.
__global__ __launch_bounds__(1024) void k_trsv(const int* __restrict__ rp, const int* __restrict__ ci,
                                               const double* __restrict__ v, int nlv, const int* __restrict__ lv_off,
                                               const int* __restrict__ lv_rows, const double* __restrict__ z,
                                               double* __restrict__ w, int upper) {
    for (int L = 0; L < nlv; L++) {
        const int e = lv_off[L + 1];
        for (int t = lv_off[L] + threadIdx.x; t < e; t += blockDim.x) {
            const int i = lv_rows[t];
            const int b = rp[i], en = rp[i + 1];
            double sum = z[i];
            if (!upper) {
                for (int k = b; k < en - 1; k++) sum -= v[k] * w[ci[k]];
                sum /= v[en - 1];
            } else {
                const double d = v[b];
                for (int k = b + 1; k < en; k++) sum -= v[k] * w[ci[k]];
                sum /= d;
            }
            w[i] = sum;
        }
        __syncthreads();
    }
}

// The same two sweeps with the factors packed by level (TriPack).  One row
// per thread; each level's row records are loaded D = 2 levels ahead into
// registers and stay in flight across the barriers, which fence LDS only: the
// results a level reads come from the LDS ring, so a level costs an LDS round
// trip, the row's arithmetic and the barrier, and the record stream (about
// 72 bytes a row) is bounded by what one compute unit can load -- measured on
// the 1.05M-row damage factors, 2,049 levels a sweep: 1.35 us a level, one
// preconditioner application 5.5 ms (11.3 ms a sweep with k_trsv).  Deeper
// prefetch (4 levels) measured no faster.  The arithmetic is k_trsv's:
// s = in; s -= v_k x_k in the factor's row order; s /= diag.
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// the loaded values are only stored here (nothing computed from them), so
// that the loads stay in flight until the row is solved
// SOR mode (lexicographic SOR sweep, SorPack): a row also carries
// a_t = (1 - omega) x_t (old x) and the products of its entries whose
// values are the sweep's old ones, computed by k_sor_gather (zero padded);
// the result is a_t + (omega s) / diag.
struct SorArgs {
    const double* a = nullptr;
    const double* prod = nullptr;
    double omega = 0;
    const int* ocnt = nullptr;       // old entries per position (past MAXP: prodov[ooff[t] ...])
    const int* ooff = nullptr;
    const double* prodov = nullptr;
    int old_first = 0;               // the old entries' products are already in the right-hand side
};

template <int MAXP, bool SOR>
struct TriRow {
    TriRec<MAXP> r;
    double in;
    double a;
    double prod[SOR ? MAXP : 1];
    int oc;
    bool ok;
};

// row t of the level ending at e (rows at or past e: ok = false, loads
// clamped to row 0); only stored, nothing computed from the loaded values,
// so that the loads stay in flight until the row is solved
template <int MAXP, bool SOR>
__device__ __forceinline__ void tri_load(TriRow<MAXP, SOR>& r, const TriPack& a, const SorArgs& so, const double* in,
                                         int t, int e) {
    r.ok = t < e;
    const int tt = r.ok ? t : 0;
    r.r = static_cast<const TriRec<MAXP>*>(a.rec)[tt];
    r.in = in[tt];
    if (SOR) {
        r.a = so.a[tt];
        r.oc = so.ocnt[tt];
#pragma unroll
        for (int k = 0; k < MAXP; k++) r.prod[k] = so.prod[k * a.n + tt];
    }
}

template <int MAXP, bool SOR>
__device__ __forceinline__ void tri_finish(const TriRow<MAXP, SOR>& row, const SorArgs& so, double s, double* ring,
                                           double* out, int t) {
    double x;
    if (SOR) {
        if (!so.old_first) {
#pragma unroll
            for (int k = 0; k < MAXP; k++) s -= row.prod[k];
            if (row.oc > MAXP) {
                const double* q = so.prodov + so.ooff[t];
                for (int k = 0; k < row.oc - MAXP; k++) s -= q[k];
            }
        }
        x = row.a + (so.omega * s) / row.r.diag;
    } else {
        x = s / row.r.diag;
    }
    ring[t & (TRI_RING - 1)] = x;
    out[row.r.orow] = x;
}

// a level whose sources all lie in the ring: unconditional LDS reads (padding
// slots hold value 0 and read the zero slot TRI_RING, s - 0 * 0 = s)
template <int MAXP, bool SOR>
__device__ __forceinline__ void tri_row_lds(const TriRow<MAXP, SOR>& row, const TriPack& a, const SorArgs& so,
                                            double* ring, double* out, int t) {
    if (!row.ok) return;
    const TriRec<MAXP>& r = row.r;
    double s = row.in;
#pragma unroll
    for (int k = 0; k < MAXP; k++) s -= r.val[k] * ring[r.src[k]];
    if (r.cnt > MAXP) {
        const int o = a.ovp[t] - MAXP;
        for (int k = MAXP; k < r.cnt; k++) s -= a.ovval[o + k] * ring[a.ovsrc[o + k]];
    }
    tri_finish(row, so, s, ring, out, t);
}

// a level that also reads results through global memory
template <int MAXP, bool SOR>
__device__ __forceinline__ void tri_row_any(const TriRow<MAXP, SOR>& row, const TriPack& a, const SorArgs& so,
                                            double* ring, double* out, int t) {
    if (!row.ok) return;
    const TriRec<MAXP>& r = row.r;
    double s = row.in;
#pragma unroll
    for (int k = 0; k < MAXP; k++)
        if (k < r.cnt) {
            const int q = r.src[k];
            const double x = q >= 0 ? ring[q] : out[-q - 1];
            s -= r.val[k] * x;
        }
    if (r.cnt > MAXP) {
        const int o = a.ovp[t] - MAXP;
        for (int k = MAXP; k < r.cnt; k++) {
            const int q = a.ovsrc[o + k];
            const double x = q >= 0 ? ring[q] : out[-q - 1];
            s -= a.ovval[o + k] * x;
        }
    }
    tri_finish(row, so, s, ring, out, t);
}

// Level offsets are read from LDS (staged TRI_OFFCAP at a time): a global
// read of them would wait behind the row prefetches in flight.
constexpr int TRI_OFFCAP = 16384;

template <int MAXP, int D, bool SOR>
__device__ __forceinline__ void tri_phase(const TriPack& a, const SorArgs& so, const double* in, double* out,
                                          double* ring, int* loff) {
    constexpr int CH = ((TRI_OFFCAP - 1 - D) / D) * D;  // levels per staged window
    constexpr int M = 0x7fffffff;
    const int tid = threadIdx.x;
    for (int q = tid; q < TRI_OFFCAP && q <= a.nlv; q += TRI_NT) loff[q] = a.off[q];
    __syncthreads();
    TriRow<MAXP, SOR> st[D];
#pragma unroll
    for (int d = 0; d < D; d++) {
        // stage d's loads issued before stage d+1's, as in the loop
        __builtin_amdgcn_sched_barrier(0);
        int b = 0, e = 0;
        if (d < a.nlv) {
            b = loff[d] & M;
            e = loff[d + 1] & M;
        }
        tri_load(st[d], a, so, in, b + tid, e);
    }
    __builtin_amdgcn_sched_barrier(0);
    int c0 = 0;
    for (int L0 = 0; L0 < a.nlv; L0 += D) {
        if (L0 - c0 >= CH) {
            // the previous level's barrier ended every read of the old window
            c0 = L0;
            for (int q = tid; q < TRI_OFFCAP && c0 + q <= a.nlv; q += TRI_NT) loff[q] = a.off[c0 + q];
            lds_barrier();
        }
#pragma unroll
        for (int d = 0; d < D; d++) {
            // keeps the scheduler from hoisting a later stage's address
            // arithmetic (and so its load waits) above this stage
            __builtin_amdgcn_sched_barrier(0);
            const int L = L0 + d;
            if (L < a.nlv) {
                const int v0 = loff[L - c0];
                const int b = v0 & M, e = loff[L + 1 - c0] & M;
                const bool glob = v0 < 0;
                if (glob)
                    tri_row_any(st[d], a, so, ring, out, b + tid);
                else
                    tri_row_lds(st[d], a, so, ring, out, b + tid);
                // rows past the first TRI_NT of a wide level: loaded here
                for (int t = b + TRI_NT + tid; t < e; t += TRI_NT) {
                    TriRow<MAXP, SOR> r;
                    tri_load(r, a, so, in, t, e);
                    if (glob)
                        tri_row_any(r, a, so, ring, out, t);
                    else
                        tri_row_lds(r, a, so, ring, out, t);
                }
                const int L2 = L + D;
                int b2 = 0, e2 = 0;
                if (L2 < a.nlv) {
                    b2 = loff[L2 - c0] & M;
                    e2 = loff[L2 + 1 - c0] & M;
                }
                tri_load(st[d], a, so, in, b2 + tid, e2);
                if (L + 1 < a.nlv && loff[L + 1 - c0] < 0)
                    __syncthreads();
                else
                    lds_barrier();
            }
        }
    }
}

template <int MAXP, int D>
__global__ __launch_bounds__(TRI_NT) void k_tri_pair(TriPack fw, TriPack bw, const double* __restrict__ zp,
                                                      double* y, double* w) {
    __shared__ double ring[TRI_RING + 1];
    __shared__ int loff[TRI_OFFCAP];
    if (threadIdx.x == 0) ring[TRI_RING] = 0.0;
    const SorArgs none;
    tri_phase<MAXP, D, false>(fw, none, zp, y, ring, loff);
    __syncthreads();
    tri_phase<MAXP, D, false>(bw, none, y, w, ring, loff);
}

// one exact lexicographic SOR sweep: in = the gathered right-hand side (f,
// or f minus the old entries' products when those come first in the rows),
// out = x (natural order)
template <int MAXP, int D>
__global__ __launch_bounds__(TRI_NT) void k_sor_sweep(TriPack p, SorArgs so, const double* __restrict__ in,
                                                       double* x) {
    __shared__ double ring[TRI_RING + 1];
    __shared__ int loff[TRI_OFFCAP];
    if (threadIdx.x == 0) ring[TRI_RING] = 0.0;
    tri_phase<MAXP, D, true>(p, so, in, x, ring, loff);
}

// before a sweep, per phase position t (row i): a_t = (1 - omega) x_i and,
// for the row's entries whose values are the sweep's old ones (x before it),
// old_first: in_t = f_i - sum of their products in row order; else in_t = f_i
// and prod[k n + t] = their products (k < ocnt_t)
__global__ void k_sor_gather(int n, int maxp, const int* __restrict__ row, const int* __restrict__ ocnt,
                             const double* __restrict__ oval, const int* __restrict__ ocol,
                             const int* __restrict__ ooff, const double* __restrict__ oovv,
                             const int* __restrict__ oovc, int old_first, const double* __restrict__ f,
                             const double* __restrict__ x, double omega, double* __restrict__ in,
                             double* __restrict__ a, double* __restrict__ prod, double* __restrict__ prodov) {
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
        const int i = row[t], c = ocnt[t];
        a[t] = (1.0 - omega) * x[i];
        double s = f[i];
        for (int k = 0; k < c; k++) {
            const bool slot = k < maxp;
            const int o = slot ? k * n + t : ooff[t] + k - maxp;
            const double q = (slot ? oval[o] : oovv[o]) * x[slot ? ocol[o] : oovc[o]];
            if (old_first) s -= q;
            else if (slot) prod[o] = q;
            else prodov[o] = q;
        }
        in[t] = s;
    }
}

__global__ void k_gather_perm(int n, const int* __restrict__ perm, const double* __restrict__ z,
                              double* __restrict__ zp) {
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) zp[t] = z[perm[t]];
}

template <int MAXP>
__global__ void k_tri_values(int n, const int* __restrict__ vmap, const int* __restrict__ dmap, int nov,
                             const int* __restrict__ ovmap, const double* __restrict__ v, TriRec<MAXP>* __restrict__ rec,
                             double* __restrict__ ovval) {
    const int s = gridDim.x * blockDim.x;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += s) {
        rec[t].diag = v[dmap[t]];
#pragma unroll
        for (int k = 0; k < MAXP; k++) {
            const int q = vmap[k * n + t];
            rec[t].val[k] = q >= 0 ? v[q] : 0.0;
        }
    }
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < nov; q += s) ovval[q] = v[ovmap[q]];
}

// ILU(0) numeric factorisation (incomplete_LU_factorization,
// linear_algebra.c:1952-1982, with row_times_row :1795-1819) of A with rows
// sorted by column, into the fixed patterns L (row i: the lower columns of A's
// row i, then i) and U (row i: i, then the upper columns).  Row i (one
// thread; the rows of a level in parallel, levels in order) walks its A
// entries in column order:
//   s = sum, k ascending over the L entries (i, k) computed so far (k < j,
//       and k < i), of L(i,k) U(k,j) where U has (k, j)
//   j < i: L(i,j) = a_ij - s;  j == i: L(i,i) = 0 + (a_ii - s)
//   j > i: U(i,j) = (a_ij - s) / L(i,i)   (a zero pivot sets *fail)
// U(i,i) = 1.  The same operations in the same order as the host ilu0, so
// the factors are bit-identical.
__device__ __forceinline__ double ilu_dot(int q0, int q1, const int* __restrict__ ciL, const double* vL,
                                          const int* __restrict__ rpU, const int* __restrict__ ciU,
                                          const double* vU, int j) {
    double s = 0.0;
    for (int q = q0; q < q1; q++) {
        const int k = ciL[q];
        int lo = rpU[k] + 1, hi = rpU[k + 1] - 1;  // U row k: diagonal first, then ascending
        while (lo <= hi) {
            const int mid = (lo + hi) >> 1;
            const int c = ciU[mid];
            if (c == j) {
                s += vL[q] * vU[mid];
                break;
            }
            if (c < j) lo = mid + 1;
            else hi = mid - 1;
        }
    }
    return s;
}

__global__ __launch_bounds__(1024) void k_ilu0(int nlv, const int* __restrict__ lv_off,
                                               const int* __restrict__ lv_rows, const int* __restrict__ rpA,
                                               const int* __restrict__ ciA, const double* __restrict__ vA,
                                               const int* __restrict__ rpL, const int* __restrict__ ciL, double* vL,
                                               const int* __restrict__ rpU, const int* __restrict__ ciU, double* vU,
                                               int* __restrict__ fail) {
    for (int L = 0; L < nlv; L++) {
        const int e = lv_off[L + 1];
        for (int t = lv_off[L] + threadIdx.x; t < e; t += blockDim.x) {
            const int i = lv_rows[t];
            const int l0 = rpL[i], ld = rpL[i + 1] - 1;
            int pl = l0, pu = rpU[i] + 1;
            vU[rpU[i]] = 1.0;
            vL[ld] = 0.0;
            for (int p = rpA[i]; p < rpA[i + 1]; p++) {
                const int j = ciA[p];
                const double a = vA[p];
                const double s = ilu_dot(l0, pl, ciL, vL, rpU, ciU, vU, j);
                if (j < i) {
                    vL[pl++] = a - s;
                } else if (j == i) {
                    vL[ld] = 0.0 + (a - s);
                } else {
                    const double d = vL[ld];
                    if (d != 0.0) vU[pu] = (a - s) / d;
                    else *fail = 1;
                    pu++;
                }
            }
        }
        __syncthreads();
    }
}

// k_ilu0 without the searches: row t of the level order (one thread, levels
// in order) takes its products' U entries from the plan, keeps its L values
// in registers, and performs k_ilu0's operations in k_ilu0's order (so the
// factors are bit-identical): per A entry m in column order, s = sum over
// the mask's slots q ascending of L(i,k_q) U(k_q,j); m < nl: L = a - s;
// the diagonal: L(i,i) = 0 + (a - s); above it: U = (a - s) / L(i,i).
// one row (position t of L's level schedule) of the planned factorisation
__device__ __forceinline__ void ilu0_planned_row(int t, const int4* __restrict__ m1, const int4* __restrict__ m2,
                                                 const int* __restrict__ mask, const int* __restrict__ pu,
                                                 const double* __restrict__ vA, double* vL, double* vU,
                                                 int* __restrict__ fail) {
    const int4 a = m1[t], b = m2[t];
    const int a0 = a.x, na = a.y, l0 = a.z, u0 = a.w, nl = b.x, hasd = b.y;
    int r = b.z, pw = u0 + 1;
    double Lr[ILU_MAXL];
#pragma unroll
    for (int q = 0; q < ILU_MAXL; q++) Lr[q] = 0.0;
    double dg = 0.0;
    vU[u0] = 1.0;
#pragma unroll
    for (int m = 0; m < ILU_MAXA; m++) {
        if (m < na) {
            const double av = vA[a0 + m];
            const int mk = mask[a0 + m];
            double s = 0.0;
#pragma unroll
            for (int q = 0; q < ILU_MAXL; q++)
                if ((mk >> q) & 1) {
                    s += Lr[q] * vU[pu[r]];
                    r++;
                }
            if (m < ILU_MAXL && m < nl) {
                Lr[m < ILU_MAXL ? m : 0] = av - s;
            } else if (hasd && m == nl) {
                dg = 0.0 + (av - s);
            } else {
                if (dg != 0.0) vU[pw] = (av - s) / dg;
                else *fail = 1;
                pw++;
            }
        }
    }
#pragma unroll
    for (int q = 0; q < ILU_MAXL; q++)
        if (q < nl) vL[l0 + q] = Lr[q];
    vL[l0 + nl] = dg;
}

__global__ __launch_bounds__(1024) void k_ilu0_planned(int nlv, const int* __restrict__ lv_off,
                                                       const int4* __restrict__ m1, const int4* __restrict__ m2,
                                                       const int* __restrict__ mask, const int* __restrict__ pu,
                                                       const double* __restrict__ vA, double* vL, double* vU,
                                                       int* __restrict__ fail) {
    for (int L = 0; L < nlv; L++) {
        const int e = lv_off[L + 1];
        for (int t = lv_off[L] + threadIdx.x; t < e; t += blockDim.x) ilu0_planned_row(t, m1, m2, mask, pu, vA, vL, vU, fail);
        __syncthreads();
    }
}

// the same for one wide level [t0, t1) on a full grid (a launch per level:
// the rows of a level depend only on earlier levels)
__global__ __launch_bounds__(TPB) void k_ilu0_planned_level(int t0, int t1, const int4* __restrict__ m1,
                                                            const int4* __restrict__ m2, const int* __restrict__ mask,
                                                            const int* __restrict__ pu, const double* __restrict__ vA,
                                                            double* vL, double* vU, int* __restrict__ fail) {
    for (int t = t0 + blockIdx.x * TPB + threadIdx.x; t < t1; t += gridDim.x * TPB)
        ilu0_planned_row(t, m1, m2, mask, pu, vA, vL, vU, fail);
}

// One level [t0, t1) of a triangular sweep on a full grid, a row per
// thread, each row's arithmetic that of PCG_preconl / PCG_preconr: forward
// (L, diagonal last) s = z_i, s -= l_ij y_j in column order, y_i = s / l_ii;
// backward (U, diagonal first) s = y_i, s -= u_ij w_j, w_i = s / u_ii.
__global__ __launch_bounds__(TPB) void k_tri_level(int t0, int t1, const int* __restrict__ rows,
                                                   const int* __restrict__ rp, const int* __restrict__ ci,
                                                   const double* __restrict__ v, const double* __restrict__ in,
                                                   double* out, int upper) {
    for (int t = t0 + blockIdx.x * TPB + threadIdx.x; t < t1; t += gridDim.x * TPB) {
        const int i = rows[t];
        const int b = rp[i], e = rp[i + 1];
        double s = in[i];
        if (!upper) {
            for (int p = b; p < e - 1; p++) s -= v[p] * out[ci[p]];
            s /= v[e - 1];
        } else {
            for (int p = b + 1; p < e; p++) s -= v[p] * out[ci[p]];
            s /= v[b];
        }
        out[i] = s;
    }
}

// out[r] = in[perm[r]] (inverse == 0) or out[perm[r]] = in[r] (inverse == 1)
__global__ void k_permute(int n, const int* __restrict__ perm, const double* __restrict__ in, double* __restrict__ out,
                          int inverse) {
    for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x) {
        if (inverse) out[perm[r]] = in[r];
        else out[r] = in[perm[r]];
    }
}

// h_l = V_l . r for l < m (classical Gram-Schmidt dots), block partials
// part[l * nblk + block]
__global__ __launch_bounds__(TPB) void k_multidot(int n, int m, const double* __restrict__ V, size_t ld,
                                                  const double* __restrict__ r, double* __restrict__ part) {
    for (int l = 0; l < m; l++) {
        const double* vl = V + (size_t)l * ld;
        double s = 0.0;
        for (int i = blockIdx.x * TPB + threadIdx.x; i < n; i += gridDim.x * TPB) s += vl[i] * r[i];
        double t = block_sum(s);
        if (threadIdx.x == 0) part[(size_t)l * gridDim.x + blockIdx.x] = t;
    }
}

// out[l] = sum of the nblk partials of dot l (fixed order)
__global__ __launch_bounds__(TPB) void k_multidot_final(int m, int nblk, const double* __restrict__ part,
                                                        double* __restrict__ out) {
    for (int l = 0; l < m; l++) {
        double s = 0.0;
        for (int i = threadIdx.x; i < nblk; i += TPB) s += part[(size_t)l * nblk + i];
        double t = block_sum(s);
        if (threadIdx.x == 0) out[l] = t;
    }
}

// y[j] = base[j] + sign * sum_{l<m} V_l[j] * c[l]   (c on the device)
__global__ void k_combine(int n, int m, const double* __restrict__ V, size_t ld, const double* __restrict__ c,
                          const double* __restrict__ base, double sign, double* __restrict__ y) {
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
        double s = 0.0;
        for (int l = 0; l < m; l++) s += V[(size_t)l * ld + j] * c[l];
        y[j] = base[j] + sign * s;
    }
}

// ------------------------------------------------- Galerkin products

// Plan sums: out[e] = the sum of entry e's listed products va[ia[t]] *
// vb[ib[t]], t in [ptr[e], ptr[e+1]), in their listed order -- started from
// 0.0 when seed[e] (sparse_zero's seed the reference adds onto), else from
// the first product.  A workgroup takes TPB consecutive entries, whose
// products are contiguous in the plan; all lanes gather and multiply them
// chunk by chunk into LDS, then each lane adds its own entry's products from
// LDS in their listed order (every sum sequential, as the reference's),
// while the dependent gathers run in parallel.
constexpr int PLAN_CHUNK = 4096;
__global__ __launch_bounds__(TPB) void k_plan_sum_staged(int ne, const int64_t* __restrict__ ptr,
                                                        const int* __restrict__ ia, const int* __restrict__ ib,
                                                        const uint8_t* __restrict__ seed,
                                                        const double* __restrict__ va, const double* __restrict__ vb,
                                                        double* __restrict__ out) {
    __shared__ double prod[PLAN_CHUNK];
    const int e0 = blockIdx.x * TPB;
    const int e = e0 + (int)threadIdx.x;
    const int elast = min(e0 + TPB, ne);
    const int64_t b0 = ptr[e0], b1 = ptr[elast];
    const bool live = e < ne;
    const int64_t my0 = live ? ptr[e] : 0, my1 = live ? ptr[e + 1] : 0;
    bool started = live && seed[e];
    double s = 0.0;
    for (int64_t c0 = b0; c0 < b1; c0 += PLAN_CHUNK) {
        const int64_t c1 = min(c0 + (int64_t)PLAN_CHUNK, b1);
        for (int64_t t = c0 + threadIdx.x; t < c1; t += TPB) prod[t - c0] = va[ia[t]] * vb[ib[t]];
        __syncthreads();
        const int64_t lo = max(my0, c0), hi = min(my1, c1);
        for (int64_t t = lo; t < hi; t++) {
            const double p = prod[t - c0];
            if (started) {
                s += p;
            } else {
                s = p;
                started = true;
            }
        }
        __syncthreads();
    }
    if (live) out[e] = s;
}

// symmetric Dirichlet treatment keeping the pattern, one row per thread:
// a constrained row becomes the unit row with b[r] = u[r]; any other row
// subtracts sum over its constrained columns c (column order, from +0.0) of
// a_rc u_c from b[r] and zeroes those entries
__global__ void k_dirichlet_rows(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                                 const uint8_t* __restrict__ dir, const double* __restrict__ u,
                                 double* __restrict__ val, double* __restrict__ b) {
    for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x) {
        const int p0 = rp[r], p1 = rp[r + 1];
        if (dir[r]) {
            for (int p = p0; p < p1; p++) val[p] = ci[p] == r ? 1.0 : 0.0;
            if (b) b[r] = u[r];
            continue;
        }
        double s = 0.0;
        bool any = false;
        for (int p = p0; p < p1; p++) {
            const int c = ci[p];
            if (dir[c]) {
                s += val[p] * u[c];
                val[p] = 0.0;
                any = true;
            }
        }
        if (b && any) b[r] = b[r] - s;
    }
}

// v = -v over the listed rows (the setup's diagonal signs)
__global__ void k_negate_rows(int m, const int* __restrict__ rows, const int* __restrict__ rp, double* __restrict__ v) {
    for (int k = blockIdx.x; k < m; k += gridDim.x) {
        const int r = rows[k];
        for (int p = rp[r] + threadIdx.x; p < rp[r + 1]; p += blockDim.x) v[p] = -v[p];
    }
}

__global__ void k_plan_scatter(int ne, const double* __restrict__ val, const int* __restrict__ pos,
                               const uint8_t* __restrict__ lone, double* __restrict__ dst, int* __restrict__ flag) {
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < ne; e += gridDim.x * blockDim.x) {
        const double v = val[e];
        const int q = pos[e];
        if (q >= 0) {
            if (v != 0.0) dst[q] = v;
            else if (lone[e]) dst[q] = 0.0;
            else atomicOr(flag, 1);
        } else if (v != 0.0) {
            atomicOr(flag, 1);
        }
    }
}

template <typename VT>
__global__ void k_row_l1(int n, const int* __restrict__ rp, const double* __restrict__ v, VT* __restrict__ out) {
    for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x) {
        VT s = 0;
        for (int p = rp[r]; p < rp[r + 1]; p++) s += fabs((VT)v[p]);
        out[r] = s;
    }
}

template <typename T>
__global__ void k_gather(int n, const int* __restrict__ idx, const T* __restrict__ src, T* __restrict__ dst) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int k = idx[i];
        dst[i] = k >= 0 ? src[k] : (T)0;
    }
}

__global__ void k_scatter(int n, const int* __restrict__ idx, const double* __restrict__ src, double* __restrict__ dst) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int k = idx[i];
        if (k >= 0) dst[k] = src[i];
    }
}

// ------------------------------------------------- reflective Newton
//
// Box-constraint bookkeeping of reflective_newton.c over the bounded index
// set idx[0..nb) with bounds lo/up (NaN: unbounded side).

// RN_set_D_J (reflective_newton.c:227-248); D = 1, J = 0 beforehand
// The reflective Newton's scaled Hessian M = D H D + diag(|g| J) on the
// device, row by row, with the host assembly's rules (RN_get_M_D's
// two products with their zero deletions, then insert_sparse of |g_i| J_i):
// over the row's positions in F = H's pattern plus the diagonal (fh: H has
// the entry, fhv its value):
//   HD: off-diagonal (i,k) present iff h_ik != 0 and d_k != 0, value
//       h_ik d_k; the diagonal always, 0.0 + (h_ii d_i or 0) (a 0.0 seed
//       without h_ii); then the zeros deleted when the row has more than one
//       entry, the diagonal 0.0 kept when nothing else remains;
//   M:  off-diagonal present iff d_i != 0 (its HD value is nonzero), value
//       d_i (h_ik d_k); the diagonal 0.0 + d_i b (b its HD value, both
//       nonzero) or a 0.0 seed; the same zero deletion; then |g_i| J_i added
//       to the diagonal (inserted when J_i != 0 and no diagonal is left).
// mpos: each F position's entry in the cached pattern of M (-1: absent);
// a position whose presence differs sets *flag (the caller assembles on the
// host); present values go to out[mpos] (mpos == nullptr: neither); outF,
// when given, receives every F position's value, 0.0 where M has no entry.
__device__ __forceinline__ double rn_hd(int i, int k, bool has, double a, const double* __restrict__ D) {
    const double b = D[k];
    const double val = (has && a != 0 && b != 0) ? a * b : 0.0;
    return k == i ? (has ? 0.0 + val : 0.0) : val;
}

// 8 lanes a row, each lane the row's entries lane, lane + 8, ...; the
// row's counts and its diagonal gathered by xor shuffles within the 8
constexpr int RN_LPR = 8;
__global__ __launch_bounds__(TPB) void k_rn_scaled_hessian(int n, const int* __restrict__ frp,
                                                           const int* __restrict__ fci,
                                                           const double* __restrict__ fhv,
                                                           const uint8_t* __restrict__ fh,
                                                           const double* __restrict__ D,
                                                           const double* __restrict__ G,
                                                           const double* __restrict__ J,
                                                           const int* __restrict__ mpos, double* __restrict__ out,
                                                           double* __restrict__ outF, int* __restrict__ flag) {
    const int lane = threadIdx.x % RN_LPR;
    const int groups = gridDim.x * (TPB / RN_LPR);
    for (int i = blockIdx.x * (TPB / RN_LPR) + threadIdx.x / RN_LPR; i < n; i += groups) {
        const int b0 = frp[i], b1 = frp[i + 1];
        const double di = D[i];
        // HD: entries present, nonzero entries, the diagonal's value
        int hd_len = 0, hd_nz = 0;
        double hd_diag = 0.0;
        for (int p = b0 + lane; p < b1; p += RN_LPR) {
            const int k = fci[p];
            const bool has = fh[p] != 0;
            const double a = has ? fhv[p] : 0.0;
            const double v = rn_hd(i, k, has, a, D);
            const bool pres = k == i || (has && a != 0 && D[k] != 0);
            if (!pres) continue;
            hd_len++;
            hd_nz += v != 0;
            if (k == i) hd_diag = v;
        }
        hd_len = group_sum<RN_LPR>(hd_len);
        hd_nz = group_sum<RN_LPR>(hd_nz);
        hd_diag = group_sum<RN_LPR>(hd_diag);  // one lane holds it, the others add 0.0
        // after HD's deletion: the diagonal kept with value hd_diag, or 0.0
        const bool hd_keep_diag = hd_len == 1 || hd_diag != 0 || hd_nz == 0;
        // M before its deletion: the diagonal always, off-diagonals iff
        // di != 0 (over HD's kept, nonzero off-diagonals)
        const double m_diag = (hd_keep_diag && di != 0 && hd_diag != 0) ? 0.0 + di * hd_diag : 0.0;
        int m_off = 0, m_nz = 0;
        for (int p = b0 + lane; p < b1; p += RN_LPR) {
            const int k = fci[p];
            if (k == i || !fh[p]) continue;
            const double v = rn_hd(i, k, true, fhv[p], D);
            if (!(fhv[p] != 0 && D[k] != 0) || v == 0 || di == 0) continue;
            m_off++;
            m_nz += di * v != 0;
        }
        m_off = group_sum<RN_LPR>(m_off);
        m_nz = group_sum<RN_LPR>(m_nz) + (m_diag != 0);
        const bool keep_diag = m_off == 0 || m_diag != 0 || m_nz == 0;
        const double add = J[i] != 0 ? fabs(G[i]) * J[i] : 0.0;
        bool bad = false;
        for (int p = b0 + lane; p < b1; p += RN_LPR) {
            const int k = fci[p];
            bool pres;
            double val;
            if (k == i) {
                pres = keep_diag || J[i] != 0;
                val = keep_diag ? (J[i] != 0 ? m_diag + add : m_diag) : add;
            } else {
                const bool has = fh[p] != 0;
                const double v = has ? rn_hd(i, k, true, fhv[p], D) : 0.0;
                const bool hd_kept = has && fhv[p] != 0 && D[k] != 0 && v != 0;
                val = di * v;
                pres = hd_kept && di != 0 && val != 0;
            }
            if (outF) outF[p] = pres ? val : 0.0;
            if (!mpos) continue;
            const int q = mpos[p];
            if (pres != (q >= 0)) bad = true;
            else if (pres) out[q] = val;
        }
        if (bad) *flag = 1;
    }
}

__global__ void k_rn_dj(int nb, const int* __restrict__ idx, const double* __restrict__ lo,
                        const double* __restrict__ up, const double* __restrict__ G, const double* __restrict__ X,
                        double* __restrict__ D, double* __restrict__ J) {
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < nb; t += gridDim.x * blockDim.x) {
        const int i = idx[t];
        if (G[i] < 0) {
            if (!isnan(up[t])) {
                D[i] = sqrt(fabs(X[i] - up[t]));
                J[i] = 1.0;
            }
        } else if (!isnan(lo[t])) {
            D[i] = sqrt(fabs(X[i] - lo[t]));
            J[i] = 1.0;
        }
    }
}

// RN_reflect (reflective_newton.c:275-292): fold X back into the box
__global__ void k_rn_reflect(int nb, const int* __restrict__ idx, const double* __restrict__ lo,
                             const double* __restrict__ up, double* __restrict__ X) {
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < nb; t += gridDim.x * blockDim.x) {
        const int i = idx[t];
        const double l = lo[t], u = up[t];
        if (!isnan(l)) {
            if (!isnan(u)) {
                const double d = 2. * (u - l);
                double whole;
                const double w = modf(fabs(X[i] - l) / d, &whole) * d;
                X[i] = l + (w < d - w ? w : d - w);
            } else {
                X[i] = l + fabs(X[i] - l);
            }
        } else if (X[i] > u) {
            X[i] = 2. * u - X[i];
        }
    }
}

// delete_small_diff (reflective_newton.c:476-484)
__global__ void k_rn_snap(int nb, const int* __restrict__ idx, const double* __restrict__ lo,
                          const double* __restrict__ up, double* __restrict__ X, double eps) {
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < nb; t += gridDim.x * blockDim.x) {
        const int i = idx[t];
        if (!isnan(up[t]) && fabs(X[i] - up[t]) < eps) X[i] = up[t];
        if (!isnan(lo[t]) && fabs(X[i] - lo[t]) < eps) X[i] = lo[t];
    }
}

// is_at_boundary (reflective_newton.c:498-513): block partial counts of
// bounded entries within 1e-30 of a bound (the caller tests the sum > 0)
__global__ __launch_bounds__(TPB) void k_rn_touch(int nb, const int* __restrict__ idx,
                                                  const double* __restrict__ lo, const double* __restrict__ up,
                                                  const double* __restrict__ X, double* __restrict__ part) {
    double c = 0.0;
    for (int t = blockIdx.x * TPB + threadIdx.x; t < nb; t += gridDim.x * TPB) {
        const int i = idx[t];
        if ((!isnan(up[t]) && fabs(X[i] - up[t]) < 1e-30) || (!isnan(lo[t]) && fabs(X[i] - lo[t]) < 1e-30)) c += 1.0;
    }
    double s = block_sum(c);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// x_i *= y_i (vector_pseudo_mult), x_i *= a (scalar_mult)
__global__ void k_mul(int n, double* __restrict__ x, const double* __restrict__ y) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) x[i] *= y[i];
}
__global__ void k_scale(int n, double* __restrict__ x, double a) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) x[i] *= a;
}

// ------------------------------------------------- adjoint gradient
//
// dL/dtheta_e = -scale_e * lam_e^T K_e(E=1) u_e for P1 plane-strain elements
// with Young's modulus theta_e (K linear in theta_e).  With the gradients
// g_i of the barycentric coordinates and strains of u and lam,
//   lam^T K_e u = area * [(l+2m)(ax cx + by dy) + l(ax dy + by cx) + m(ay + bx)(cy + dx)]
// where (ax, ay) = grad lam_x, (bx, by) = grad lam_y, (cx, cy) = grad u_x,
// (dx, dy) = grad u_y and l, m are the Lame constants of E = 1.  u and lam
// are full vectors (constrained entries: prescribed values in u, zero in
// lam); blocked != 0: [x of all nodes; y of all nodes], else interleaved.
__global__ void k_elast_grad(int ne, const int* __restrict__ tri, const double* __restrict__ xy, int nn,
                             const double* __restrict__ u, const double* __restrict__ lam, double lame_l,
                             double lame_m, const double* __restrict__ scale, int blocked,
                             double* __restrict__ grad) {
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < ne; e += gridDim.x * blockDim.x) {
        const int n0 = tri[3 * e], n1 = tri[3 * e + 1], n2 = tri[3 * e + 2];
        const double x0 = xy[2 * n0], y0 = xy[2 * n0 + 1];
        const double d1x = xy[2 * n1] - x0, d1y = xy[2 * n1 + 1] - y0;
        const double d2x = xy[2 * n2] - x0, d2y = xy[2 * n2 + 1] - y0;
        const double det = d1x * d2y - d1y * d2x;
        const double area = 0.5 * fabs(det);
        const double g1x = d2y / det, g1y = -d2x / det;
        const double g2x = -d1y / det, g2y = d1x / det;
        const double g0x = -(g1x + g2x), g0y = -(g1y + g2y);
        const int ix0 = blocked ? n0 : 2 * n0, ix1 = blocked ? n1 : 2 * n1, ix2 = blocked ? n2 : 2 * n2;
        const int iy0 = blocked ? n0 + nn : 2 * n0 + 1, iy1 = blocked ? n1 + nn : 2 * n1 + 1,
                  iy2 = blocked ? n2 + nn : 2 * n2 + 1;
        const double ux0 = u[ix0], ux1 = u[ix1], ux2 = u[ix2], uy0 = u[iy0], uy1 = u[iy1], uy2 = u[iy2];
        const double lx0 = lam[ix0], lx1 = lam[ix1], lx2 = lam[ix2], ly0 = lam[iy0], ly1 = lam[iy1],
                     ly2 = lam[iy2];
        const double ax = lx0 * g0x + lx1 * g1x + lx2 * g2x, ay = lx0 * g0y + lx1 * g1y + lx2 * g2y;
        const double bx = ly0 * g0x + ly1 * g1x + ly2 * g2x, by = ly0 * g0y + ly1 * g1y + ly2 * g2y;
        const double cx = ux0 * g0x + ux1 * g1x + ux2 * g2x, cy = ux0 * g0y + ux1 * g1y + ux2 * g2y;
        const double dx = uy0 * g0x + uy1 * g1x + uy2 * g2x, dy = uy0 * g0y + uy1 * g1y + uy2 * g2y;
        const double form = (lame_l + 2.0 * lame_m) * (ax * cx + by * dy) + lame_l * (ax * dy + by * cx) +
                            lame_m * (ay + bx) * (cy + dx);
        grad[e] = -(scale ? scale[e] : 1.0) * area * form;
    }
}

// -------------------------------------------------- batched Krylov
//
// A batch of independent systems stored as one block-diagonal system.  Rows
// are split into chunks that never straddle two systems: chunk c covers rows
// [crow[c], crow[c+1]) of system cinst[c]; system k owns chunks
// [ichunk[k], ichunk[k+1]).  Per-system scalars live in device arrays.

// part[c] = sum over chunk c of a_i b_i (fixed tree)
__global__ __launch_bounds__(TPB) void k_seg_dot(int nchunk, const int* __restrict__ crow,
                                                 const double* __restrict__ a, const double* __restrict__ b,
                                                 double* __restrict__ part) {
    for (int c = blockIdx.x; c < nchunk; c += gridDim.x) {
        double s = 0.0;
        for (int i = crow[c] + threadIdx.x; i < crow[c + 1]; i += TPB) s += a[i] * b[i];
        double t = block_sum(s);
        if (threadIdx.x == 0) part[c] = t;
        __syncthreads();
    }
}

// out[k] = sum of system k's chunk partials (shift != 0: old[k] = out[k] first)
__global__ __launch_bounds__(TPB) void k_seg_sum(int ninst, const int* __restrict__ ichunk,
                                                 const double* __restrict__ part, double* __restrict__ out,
                                                 double* __restrict__ old) {
    for (int k = blockIdx.x; k < ninst; k += gridDim.x) {
        double s = 0.0;
        for (int c = ichunk[k] + threadIdx.x; c < ichunk[k + 1]; c += TPB) s += part[c];
        double t = block_sum(s);
        if (threadIdx.x == 0) {
            if (old) old[k] = out[k];
            out[k] = t;
        }
        __syncthreads();
    }
}

// P = Z + (s/s_old) P per system
__global__ __launch_bounds__(TPB) void k_bpcg_p(int nchunk, const int* __restrict__ crow,
                                                const int* __restrict__ cinst, double* __restrict__ P,
                                                const double* __restrict__ Z, const double* __restrict__ s,
                                                const double* __restrict__ sold, const double* __restrict__ active) {
    for (int c = blockIdx.x; c < nchunk; c += gridDim.x) {
        const int k = cinst[c];
        if (active[k] == 0.0) continue;  // an inactive system keeps its P
        const double beta = s[k] / sold[k];
        for (int i = crow[c] + threadIdx.x; i < crow[c + 1]; i += TPB) P[i] = Z[i] + beta * P[i];
    }
}

// X += a P, R -= a Q with a = s/pq for the active systems
__global__ __launch_bounds__(TPB) void k_bpcg_xr(int nchunk, const int* __restrict__ crow,
                                                 const int* __restrict__ cinst, double* __restrict__ X,
                                                 double* __restrict__ R, const double* __restrict__ P,
                                                 const double* __restrict__ Q, const double* __restrict__ s,
                                                 const double* __restrict__ pq, const double* __restrict__ active) {
    for (int c = blockIdx.x; c < nchunk; c += gridDim.x) {
        const int k = cinst[c];
        if (active[k] == 0.0) continue;  // its X and R stay as they are
        const double a = s[k] / pq[k];
        for (int i = crow[c] + threadIdx.x; i < crow[c + 1]; i += TPB) {
            X[i] += a * P[i];
            R[i] -= a * Q[i];
        }
    }
}

// after an iteration: count it for active systems, then stay active while
// |R| > tol |B| and the iteration limit is not reached
__global__ void k_bpcg_check(int ninst, const double* __restrict__ rr, const double* __restrict__ hb, double tol,
                             int max_iter, double* __restrict__ active, double* __restrict__ iters, int count) {
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < ninst; k += gridDim.x * blockDim.x) {
        if (active[k] == 0.0) continue;
        if (count) iters[k] += 1.0;
        active[k] = (sqrt(rr[k]) > tol * hb[k] && iters[k] < (double)max_iter) ? 1.0 : 0.0;
    }
}

// ------------------------------------------------------ coarse solve

// x = inv * b over the dense rows (gathered through idx), trivial rows
// x_i = b_i / diag_i.  One wave per dense row.
template <typename XT>
__global__ void k_coarse_gather(int nd, const int* __restrict__ didx, const XT* __restrict__ b,
                                double* __restrict__ bd) {
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < nd; c += gridDim.x * blockDim.x) bd[c] = (double)b[didx[c]];
}

// x = inv * bd over the dense rows (bd: the right-hand side gathered to the
// dense numbering), trivial rows x_i = b_i / diag_i.  One wave per dense
// row; the row is streamed with nontemporal 16-byte loads (nd even) into
// two accumulators per lane, summed in a fixed order.
template <typename XT>
__global__ __launch_bounds__(TPB) void k_coarse(int nd, const int* __restrict__ didx, const double* __restrict__ inv,
                                                int nt, const int* __restrict__ tidx, const double* __restrict__ tdiag,
                                                const XT* __restrict__ b, XT* __restrict__ x,
                                                const double* __restrict__ bd) {
    const int wave = (blockIdx.x * TPB + threadIdx.x) / WAVE;
    const int lane = threadIdx.x % WAVE;
    const int nw = gridDim.x * TPB / WAVE;
    for (int row = wave; row < nd; row += nw) {
        const double* a = inv + (size_t)row * nd;
        double s0 = 0.0, s1 = 0.0;
        if ((nd & 1) == 0) {
            const int h = nd / 2;
#pragma unroll 4
            for (int c = lane; c < h; c += WAVE) {
                s0 += ld_stream<true>(a + 2 * c) * bd[2 * c];
                s1 += ld_stream<true>(a + 2 * c + 1) * bd[2 * c + 1];
            }
        } else {
#pragma unroll 4
            for (int c = lane; c < nd; c += WAVE) s0 += ld_stream<true>(a + c) * bd[c];
        }
        double s = wave_sum(s0 + s1);
        if (lane == 0) x[didx[row]] = (XT)s;
    }
    for (int t = blockIdx.x * TPB + threadIdx.x; t < nt; t += gridDim.x * TPB) {
        int i = tidx[t];
        x[i] = (XT)((double)b[i] / tdiag[t]);
    }
}

typedef float f32x4 __attribute__((ext_vector_type(4)));

// Coarse solve with the symmetric inverse held as its upper triangle of
// CT x CT tiles of TT values (tile t = (ti[t], tj[t]), ti <= tj, row-major,
// zero padded), products and sums in FP64.  k_csym_tiles: one workgroup per stored
// tile (I, J) writes the tile's row sums against bd_J into part[I][J] and,
// off the diagonal, its column sums against bd_I into part[J][I];
// k_csym_reduce sums part[I][0..nct) in a fixed order into the dense rows of
// block I.  Every part[I][K] has exactly one writer.
template <typename TT>
__global__ __launch_bounds__(TPB) void k_csym_tiles(int nct, const int* __restrict__ ti, const int* __restrict__ tj,
                                                    const TT* __restrict__ tiles, const double* __restrict__ bd,
                                                    double* __restrict__ part) {
    __shared__ TT T[CT][CT + 1];
    __shared__ double bi[CT], bj[CT];
    const int t = blockIdx.x;
    const int I = ti[t], J = tj[t];
    const TT* src = tiles + (size_t)t * CT * CT;
    if constexpr (sizeof(TT) == 4) {
        const f32x4* s4 = reinterpret_cast<const f32x4*>(src);
#pragma unroll
        for (int u = 0; u < CT * CT / 4 / TPB; u++) {
            const int q = threadIdx.x + u * TPB;
            const f32x4 v = __builtin_nontemporal_load(s4 + q);
            const int r = (4 * q) / CT, c = (4 * q) % CT;
            T[r][c] = v.x;
            T[r][c + 1] = v.y;
            T[r][c + 2] = v.z;
            T[r][c + 3] = v.w;
        }
    } else {
#pragma unroll
        for (int u = 0; u < CT * CT / TPB; u++) {
            const int q = threadIdx.x + u * TPB;
            T[q / CT][q % CT] = __builtin_nontemporal_load(src + q);
        }
    }
    if (threadIdx.x < CT) bi[threadIdx.x] = bd[I * CT + threadIdx.x];
    else if (threadIdx.x < 2 * CT) bj[threadIdx.x - CT] = bd[J * CT + threadIdx.x - CT];
    __syncthreads();
    if (threadIdx.x < CT) {
        const int r = threadIdx.x;
        double s = 0.0;
        for (int c = 0; c < CT; c++) s += (double)T[r][c] * bj[c];
        part[((size_t)I * nct + J) * CT + r] = s;
    } else if (threadIdx.x < 2 * CT && J != I) {
        const int c = threadIdx.x - CT;
        double s = 0.0;
        for (int r = 0; r < CT; r++) s += (double)T[r][c] * bi[r];
        part[((size_t)J * nct + I) * CT + c] = s;
    }
}

template <typename XT>
__global__ __launch_bounds__(TPB) void k_csym_reduce(int nd, int nct, const double* __restrict__ part,
                                                     const int* __restrict__ didx, int nt, const int* __restrict__ tidx,
                                                     const double* __restrict__ tdiag, const XT* __restrict__ b,
                                                     XT* __restrict__ x) {
    constexpr int Q = TPB / CT;  // lanes per dense row, each summing every Q-th tile column
    __shared__ double acc[Q][CT];
    if (blockIdx.x < nct) {
        const int I = blockIdx.x, r = threadIdx.x % CT, q = threadIdx.x / CT;
        double s = 0.0;
        for (int K = q; K < nct; K += Q) s += part[((size_t)I * nct + K) * CT + r];
        acc[q][r] = s;
        __syncthreads();
        const int row = I * CT + r;
        if (q == 0 && row < nd) {
            double y = acc[0][r];
            for (int k = 1; k < Q; k++) y += acc[k][r];
            x[didx[row]] = (XT)y;
        }
    }
    for (int t = blockIdx.x * TPB + threadIdx.x; t < nt; t += gridDim.x * TPB) {
        int i = tidx[t];
        x[i] = (XT)((double)b[i] / tdiag[t]);
    }
}

// tiles of the upper triangle: entry (row, col) of tile t takes
// inv[min(row, col)][max(row, col)], so that the FP32 copy is exactly symmetric
template <typename TT>
__global__ void k_inv_tiles(int nd, int ntiles, const int* __restrict__ ti, const int* __restrict__ tj,
                            const double* __restrict__ inv, TT* __restrict__ tiles) {
    const size_t total = (size_t)ntiles * CT * CT;
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < total; k += (size_t)gridDim.x * blockDim.x) {
        const int t = (int)(k / (CT * CT)), e = (int)(k % (CT * CT));
        const int row = ti[t] * CT + e / CT, col = tj[t] * CT + e % CT;
        TT v = (TT)0;
        if (row < nd && col < nd) v = (TT)inv[(size_t)std::min(row, col) * nd + std::max(row, col)];
        tiles[k] = v;
    }
}

// Block-diagonal coarse solve of a batch: dense row g belongs to block
// rblk[g]; block b holds rows [off[b], off[b+1]) of didx and the inverse at
// inv + moff[b] (row-major, nd_b x nd_b).  One wave per dense row.
template <typename XT>
__global__ __launch_bounds__(TPB) void k_bcoarse(int nd, const int* __restrict__ rblk, const int* __restrict__ off,
                                                 const size_t* __restrict__ moff, const int* __restrict__ didx,
                                                 const double* __restrict__ inv, int nt, const int* __restrict__ tidx,
                                                 const double* __restrict__ tdiag, const XT* __restrict__ b,
                                                 XT* __restrict__ x, const double* __restrict__ active) {
    const int wave = (blockIdx.x * TPB + threadIdx.x) / WAVE;
    const int lane = threadIdx.x % WAVE;
    const int nw = gridDim.x * TPB / WAVE;
    for (int g = wave; g < nd; g += nw) {
        const int bb = rblk[g];
        if (active && active[bb] == 0.0) continue;  // a converged system (block bb) of a batch
        const int o = off[bb], m = off[bb + 1] - o;
        const double* a = inv + moff[bb] + (size_t)(g - o) * m;
        double s = 0.0;
        for (int c = lane; c < m; c += WAVE) s += ld_stream<true>(a + c) * (double)b[didx[o + c]];
        s = wave_sum(s);
        if (lane == 0) x[didx[g]] = (XT)s;
    }
    for (int t = blockIdx.x * TPB + threadIdx.x; t < nt; t += gridDim.x * TPB) {
        int i = tidx[t];
        x[i] = (XT)((double)b[i] / tdiag[t]);
    }
}

// Gauss-Jordan step k on every block with more than k rows (one wave per
// row); the wave of row k stores the block's pivot reciprocal
__global__ __launch_bounds__(TPB) void k_bgj_update(int nd, int k, const int* __restrict__ rblk,
                                                    const int* __restrict__ off, const size_t* __restrict__ moff,
                                                    double* __restrict__ M, double* __restrict__ piv) {
    const int wave = (blockIdx.x * TPB + threadIdx.x) / WAVE;
    const int lane = threadIdx.x % WAVE;
    const int nw = gridDim.x * TPB / WAVE;
    for (int g = wave; g < nd; g += nw) {
        const int bb = rblk[g];
        const int o = off[bb], m = off[bb + 1] - o;
        if (k >= m) continue;
        double* A = M + moff[bb];
        const int i = g - o;
        const double p = 1.0 / A[(size_t)k * m + k];
        if (i == k) {
            if (lane == 0) piv[bb] = p;
            continue;
        }
        double* row = A + (size_t)i * m;
        const double* rowk = A + (size_t)k * m;
        const double f = row[k] * p;
        for (int j = lane; j < m; j += WAVE)
            if (j != k) row[j] -= f * rowk[j];
    }
}

__global__ __launch_bounds__(TPB) void k_bgj_pivot(int nd, int k, const int* __restrict__ rblk,
                                                   const int* __restrict__ off, const size_t* __restrict__ moff,
                                                   double* __restrict__ M, const double* __restrict__ piv) {
    for (int g = blockIdx.x * TPB + threadIdx.x; g < nd; g += gridDim.x * TPB) {
        const int bb = rblk[g];
        const int o = off[bb], m = off[bb + 1] - o;
        if (k >= m) continue;
        double* A = M + moff[bb];
        const int t = g - o;
        const double p = piv[bb];
        if (t == k) {
            A[(size_t)k * m + k] = p;
        } else {
            A[(size_t)k * m + t] *= p;
            A[(size_t)t * m + k] *= -p;
        }
    }
}

// ------------------------------------------- coarse inverse (Gauss-Jordan)
//
// In-place Gauss-Jordan inversion of the dense SPD coarse matrix, no
// pivoting (the reference's LU_factorzation does not pivot either).  Step k:
//   a_ij -= a_ik a_kj / a_kk   (i,j != k)   -- k_gj_update, reads row/col k
//   a_kj /= a_kk, a_ik *= -1/a_kk, a_kk = 1/a_kk -- k_gj_pivot
__global__ __launch_bounds__(TPB) void k_gj_update(int N, int k, double* __restrict__ M, double* __restrict__ piv) {
    const double p = 1.0 / M[(size_t)k * N + k];
    if (blockIdx.x == 0 && threadIdx.x == 0) *piv = p;
    const double* rowk = M + (size_t)k * N;
    for (int i = blockIdx.x; i < N; i += gridDim.x) {
        if (i == k) continue;
        double* row = M + (size_t)i * N;
        const double f = row[k] * p;
        for (int j = threadIdx.x; j < N; j += TPB)
            if (j != k) row[j] -= f * rowk[j];
    }
}

__global__ __launch_bounds__(TPB) void k_gj_pivot(int N, int k, double* __restrict__ M, const double* __restrict__ piv) {
    const double p = *piv;
    for (int t = blockIdx.x * TPB + threadIdx.x; t < N; t += gridDim.x * TPB) {
        if (t == k) {
            M[(size_t)k * N + k] = p;
            continue;
        }
        M[(size_t)k * N + t] *= p;
        M[(size_t)t * N + k] *= -p;
    }
}

// ------------------------------- blocked Gauss-Jordan (FP64 MFMA update)
//
// The same inversion by blocks of GJB = 64 pivots (no pivoting).  For the
// block K = [k0, k0 + nk) and the others I, J:
//   D = M[K,K]^-1                        k_bgj_diag (one workgroup, in LDS)
//   M[K,J] = D M[K,J]                    k_bgj_panels (also copied to Brow)
//   Acol = M[I,K]; M[I,K] = -M[I,K] D    k_bgj_panels
//   M[I,J] -= Acol M[K,J]                k_bgj_gemm (64 x 64 tiles, FP64 MFMA)
//   M[K,K] = D
// With GJB = 1 these are the scalar steps above.
constexpr int GJB = 64;

__global__ __launch_bounds__(1024) void k_bgj_diag(int N, int k0, int nk, double* __restrict__ M,
                                                   double* __restrict__ Dinv) {
    __shared__ double D[GJB][GJB + 1];
    __shared__ double piv;
    const int tid = threadIdx.x;
    for (int t = tid; t < GJB * GJB; t += blockDim.x) {
        const int i = t / GJB, j = t % GJB;
        D[i][j] = (i < nk && j < nk) ? M[(size_t)(k0 + i) * N + k0 + j] : 0.0;
    }
    __syncthreads();
    for (int k = 0; k < nk; k++) {
        if (tid == 0) piv = 1.0 / D[k][k];
        __syncthreads();
        const double p = piv;
        for (int t = tid; t < nk * nk; t += blockDim.x) {
            const int i = t / nk, j = t % nk;
            if (i != k && j != k) D[i][j] -= D[i][k] * p * D[k][j];
        }
        __syncthreads();
        for (int t = tid; t < nk; t += blockDim.x) {
            if (t == k) continue;
            D[k][t] *= p;
            D[t][k] *= -p;
        }
        __syncthreads();
        if (tid == 0) D[k][k] = p;
        __syncthreads();
    }
    for (int t = tid; t < GJB * GJB; t += blockDim.x) {
        const int i = t / GJB, j = t % GJB;
        Dinv[t] = D[i][j];  // zero outside nk x nk
        if (i < nk && j < nk) M[(size_t)(k0 + i) * N + k0 + j] = D[i][j];
    }
}

// one thread per index t outside the block (64 threads a workgroup): the
// new row panel's column t (also into Brow, 64 x N, rows beyond nk zero) and
// the new column panel's row t (the old one into Acol, N x 64, columns
// beyond nk zero); each thread's 64 inputs are staged in LDS
typedef double f64x4 __attribute__((ext_vector_type(4)));

// M[I,J] -= Acol[I,:] Brow[:,J] over the 64 x 64 tiles outside block row and
// column kb.  4 waves a tile, each a 32 x 32 quadrant of 2 x 2 MFMA blocks
// (v_mfma_f64_16x16x4f64: lane l holds A[l%16][l/16], B[l/16][l%16] and
// D[4 v + l/16][l%16], v = 0..3 -- one output row per 16-lane group and
// register, unlike the FP32 16x16 form's 4 consecutive rows per group).
__global__ __launch_bounds__(256) void k_bgj_gemm(int N, int kb, int nt, double* __restrict__ M,
                                                  const double* __restrict__ Acol, const double* __restrict__ Brow) {
    __shared__ double As[GJB][GJB + 1];  // [row in tile][k]
    __shared__ double Bs[GJB][GJB + 1];  // [k][column in tile]
    const int tile = blockIdx.x;
    int ti = tile / (nt - 1), tj = tile % (nt - 1);
    if (ti >= kb) ti++;  // skip the pivot block's row and column of tiles
    if (tj >= kb) tj++;
    const int i0 = ti * GJB, j0 = tj * GJB;
    const int tid = threadIdx.x;
    const int w = tid / 64, l = tid % 64;
    const int qi = (w / 2) * 32, qj = (w % 2) * 32;
    // the lane's 16 entries of the M tile, loaded first so that their
    // latency overlaps the staging and the products
    double mv[2][2][4];
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++)
            for (int v = 0; v < 4; v++) {
                const int i = i0 + qi + a * 16 + 4 * v + l / 16;
                const int j = j0 + qj + b * 16 + l % 16;
                mv[a][b][v] = (i < N && j < N) ? M[(size_t)i * N + j] : 0.0;
            }
    for (int t = tid; t < GJB * GJB; t += 256) {
        const int r = t / GJB, c = t % GJB;
        As[r][c] = (i0 + r < N) ? Acol[(size_t)(i0 + r) * GJB + c] : 0.0;
        Bs[r][c] = (j0 + c < N) ? Brow[(size_t)r * N + j0 + c] : 0.0;
    }
    __syncthreads();
    f64x4 acc[2][2];
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++) acc[a][b] = (f64x4){0.0, 0.0, 0.0, 0.0};
    for (int k = 0; k < GJB; k += 4) {
        double av[2], bv[2];
        for (int a = 0; a < 2; a++) av[a] = As[qi + a * 16 + l % 16][k + l / 16];
        for (int b = 0; b < 2; b++) bv[b] = Bs[k + l / 16][qj + b * 16 + l % 16];
        for (int a = 0; a < 2; a++)
            for (int b = 0; b < 2; b++) acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[a], bv[b], acc[a][b], 0, 0, 0);
    }
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++)
            for (int v = 0; v < 4; v++) {
                const int i = i0 + qi + a * 16 + 4 * v + l / 16;
                const int j = j0 + qj + b * 16 + l % 16;
                if (i < N && j < N) M[(size_t)i * N + j] = mv[a][b][v] - acc[a][b][v];
            }
}

// The two panels of block step kb as 64 x 64 MFMA tile products (grid: nt
// row-panel tiles, then nt column-panel tiles; the pivot block's own tile
// writes zeros to its panel buffer and nothing else):
//   row panel, columns J of tile tj:  M[K,J] = D M[K,J], copied to Brow;
//   column panel, rows I of tile ti:  Acol = M[I,K], M[I,K] = -M[I,K] D.
// D = Dinv on its nk x nk corner, zero elsewhere.  The tiles touch disjoint
// parts of M (row panel: rows K, columns outside K; column panel: rows
// outside K, columns K).
__global__ __launch_bounds__(256) void k_bgj_panels(int N, int kb, int nt, int nk, double* __restrict__ M,
                                                    const double* __restrict__ Dinv, double* __restrict__ Acol,
                                                    double* __restrict__ Brow) {
    __shared__ double As[GJB][GJB + 1];  // [row][k]
    __shared__ double Bs[GJB][GJB + 1];  // [k][column]
    const int tid = threadIdx.x;
    const bool rowp = (int)blockIdx.x < nt;
    const int tb = rowp ? (int)blockIdx.x : (int)blockIdx.x - nt;
    const int k0 = kb * GJB, t0 = tb * GJB;
    if (tb == kb) {
        for (int t = tid; t < GJB * GJB; t += 256) {
            const int r = t / GJB, c = t % GJB;
            if (rowp && t0 + c < N) Brow[(size_t)r * N + t0 + c] = 0.0;
            if (!rowp && t0 + r < N) Acol[(size_t)(t0 + r) * GJB + c] = 0.0;
        }
        return;
    }
    for (int t = tid; t < GJB * GJB; t += 256) {
        const int r = t / GJB, c = t % GJB;
        if (rowp) {
            As[r][c] = (r < nk && c < nk) ? Dinv[r * GJB + c] : 0.0;
            Bs[r][c] = (r < nk && t0 + c < N) ? M[(size_t)(k0 + r) * N + t0 + c] : 0.0;
        } else {
            const double v = (t0 + r < N && c < nk) ? M[(size_t)(t0 + r) * N + k0 + c] : 0.0;
            As[r][c] = v;
            if (t0 + r < N) Acol[(size_t)(t0 + r) * GJB + c] = v;
            Bs[r][c] = (r < nk && c < nk) ? Dinv[r * GJB + c] : 0.0;
        }
    }
    __syncthreads();
    const int w = tid / 64, l = tid % 64;
    const int qi = (w / 2) * 32, qj = (w % 2) * 32;
    f64x4 acc[2][2];
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++) acc[a][b] = (f64x4){0.0, 0.0, 0.0, 0.0};
    for (int k = 0; k < GJB; k += 4) {
        double av[2], bv[2];
        for (int a = 0; a < 2; a++) av[a] = As[qi + a * 16 + l % 16][k + l / 16];
        for (int b = 0; b < 2; b++) bv[b] = Bs[k + l / 16][qj + b * 16 + l % 16];
        for (int a = 0; a < 2; a++)
            for (int b = 0; b < 2; b++) acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[a], bv[b], acc[a][b], 0, 0, 0);
    }
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++)
            for (int v = 0; v < 4; v++) {
                const int r = qi + a * 16 + 4 * v + l / 16;  // row of the tile product
                const int c = qj + b * 16 + l % 16;          // its column
                const double x = acc[a][b][v];
                if (rowp) {
                    if (t0 + c < N) {
                        Brow[(size_t)r * N + t0 + c] = r < nk ? x : 0.0;
                        if (r < nk) M[(size_t)(k0 + r) * N + t0 + c] = x;
                    }
                } else if (t0 + r < N && c < nk) {
                    M[(size_t)(t0 + r) * N + k0 + c] = -x;
                }
            }
}

// ---------------------------------------------------------- launchers

static inline int grid_for(int n, int cap = 4096) {
    int g = (n + TPB - 1) / TPB;
    if (g < 1) g = 1;
    return g > cap ? cap : g;
}

#define LINA_LPR_DISPATCH(LPR_VAL, ...)                    \
    switch (LPR_VAL) {                                     \
        case 1: { constexpr int L = 1; __VA_ARGS__; break; }   \
        case 4: { constexpr int L = 4; __VA_ARGS__; break; }   \
        case 8: { constexpr int L = 8; __VA_ARGS__; break; }   \
        case 16: { constexpr int L = 16; __VA_ARGS__; break; } \
        default: { constexpr int L = 32; __VA_ARGS__; break; } \
    }

// workgroups of a row-kernel launch: one per row block, or a persistent grid
// of at most LINA_SPMV_GRID workgroups (rounded down to a multiple of 8; read
// from the environment once, 0: off)
static int row_grid(int nblk) {
    static const int cap = [] {
        const char* e = std::getenv("LINA_SPMV_GRID");
        const int v = e ? std::atoi(e) : 0;
        return v > 0 ? (v / 8) * 8 : 0;
    }();
    return (cap > 0 && nblk > cap) ? cap : nblk;
}

template <typename VT, typename XT>
int spmv(hipStream_t st, const DMat<VT>& A, const XT* x, const XT* b, XT* y, int mode, double* part, const XT* w,
         const Epi<VT, XT>* epi) {
    if (A.nblk == 0) return 0;
    dim3 g(row_grid(A.nblk)), t(TPB);
    const Epi<VT, XT> e = epi ? *epi : Epi<VT, XT>{};
#define LINA_SPMV_MODE(M)                                                                                       \
    LINA_LPR_DISPATCH(A.lpr, hipLaunchKernelGGL((k_spmv<VT, XT, M, L>), g, t, 0, st, A.nblk, A.bk, A.rp, A.ci, \
                                                A.cs, A.cb, A.v, x, b, y, part, w, e, A.bspec, A.active))
    switch (mode) {
        case MODE_SET: LINA_SPMV_MODE(MODE_SET); break;
        case MODE_OUTER: LINA_SPMV_MODE(MODE_OUTER); break;
        case MODE_RES: LINA_SPMV_MODE(MODE_RES); break;
        case MODE_ADD: LINA_SPMV_MODE(MODE_ADD); break;
        case MODE_ADDB: LINA_SPMV_MODE(MODE_ADDB); break;
        case MODE_SETX: LINA_SPMV_MODE(MODE_SETX); break;
        default: return -1;
    }
#undef LINA_SPMV_MODE
    return A.nblk;
}
template int spmv<double, double>(hipStream_t, const DMat<double>&, const double*, const double*, double*, int,
                                  double*, const double*, const Epi<double, double>*);

// the outer CG's y = A x (MODE_OUTER, FP64) with HIP events recorded by the
// launch itself at the kernel's start and end (hipExtLaunchKernel), so that
// their interval is the kernel's own duration
int spmv_outer_timed(hipStream_t st, const DMat<double>& A, const double* x, double* y, double* part,
                     const double* w, hipEvent_t e0, hipEvent_t e1) {
    if (A.nblk == 0) return 0;
    dim3 g(row_grid(A.nblk)), t(TPB);
    const Epi<double, double> e{};
    LINA_LPR_DISPATCH(A.lpr, hipExtLaunchKernelGGL((k_spmv<double, double, MODE_OUTER, L>), g, t, 0, st, e0, e1, 0,
                                                   A.nblk, A.bk, A.rp, A.ci, A.cs, A.cb, A.v, x,
                                                   (const double*)nullptr, y, part, w, e, A.bspec, A.active))
    return A.nblk;
}
template int spmv<float, float>(hipStream_t, const DMat<float>&, const float*, const float*, float*, int, double*,
                                const float*, const Epi<float, float>*);

template <typename XT>
int dot_part(hipStream_t st, int n, const XT* a, const XT* b, double* part) {
    int g = grid_for(n, DOT_PART_MAX);
    hipLaunchKernelGGL((k_dot_part<XT>), dim3(g), dim3(TPB), 0, st, n, a, b, part);
    return g;
}
template int dot_part<double>(hipStream_t, int, const double*, const double*, double*);
template int dot_part<float>(hipStream_t, int, const float*, const float*, double*);

void finalize(hipStream_t st, int np, const double* part, double* scal, int slot, int op) {
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(FTPB), 0, st, np, part, scal, slot, op);
}

void finalize_stop(hipStream_t st, int np, const double* part, double* scal, double tol, int relres, int max_iter) {
    hipLaunchKernelGGL(k_finalize_stop, dim3(1), dim3(FTPB), 0, st, np, part, scal, tol, relres, max_iter);
}

template <typename XT>
int precon_in(hipStream_t st, int n, const double* F, double* part, double* X, XT* xf, XT* ff, const uint32_t* isb,
              const double* isa) {
    int g = grid_for(n, 1024);
    hipLaunchKernelGGL((k_precon_in<XT>), dim3(g), dim3(TPB), 0, st, n, F, part, X, xf, ff, isb, isa);
    return g;
}
template int precon_in<double>(hipStream_t, int, const double*, double*, double*, double*, double*, const uint32_t*,
                               const double*);
template int precon_in<float>(hipStream_t, int, const double*, double*, double*, float*, float*, const uint32_t*,
                              const double*);

template <typename XT>
int precon_out(hipStream_t st, int n, const XT* xs, double* Z, const double* F, const uint32_t* isb,
               const double* isa, const double* scal, int slot, const double* w, double* part) {
    int g = grid_for(n, 1024);
    hipLaunchKernelGGL((k_precon_out<XT>), dim3(g), dim3(TPB), 0, st, n, xs, Z, F, isb, isa, scal, slot, w, part);
    return g;
}
template int precon_out<double>(hipStream_t, int, const double*, double*, const double*, const uint32_t*,
                                const double*, const double*, int, const double*, double*);
template int precon_out<float>(hipStream_t, int, const float*, double*, const double*, const uint32_t*,
                               const double*, const double*, int, const double*, double*);

template <typename T>
void fill(hipStream_t st, int n, T* x, T val) {
    if (n <= 0) return;
    hipLaunchKernelGGL((k_fill<T>), dim3(grid_for(n)), dim3(TPB), 0, st, n, x, val);
}
template void fill<double>(hipStream_t, int, double*, double);
template void fill<float>(hipStream_t, int, float*, float);
template void fill<int>(hipStream_t, int, int*, int);

template <typename TI, typename TO>
void convert(hipStream_t st, int n, const TI* a, TO* b) {
    if (n <= 0) return;
    hipLaunchKernelGGL((k_convert<TI, TO>), dim3(grid_for(n)), dim3(TPB), 0, st, n, a, b);
}
template void convert<double, float>(hipStream_t, int, const double*, float*);
template void convert<float, double>(hipStream_t, int, const float*, double*);

template <typename T>
void copy(hipStream_t st, int n, const T* a, T* b) {
    if (n <= 0) return;
    hipLaunchKernelGGL((k_copy<T>), dim3(grid_for(n)), dim3(TPB), 0, st, n, a, b);
}
template void copy<double>(hipStream_t, int, const double*, double*);
template void copy<float>(hipStream_t, int, const float*, float*);

void signs(hipStream_t st, int n, const double* s, const double* x, double* y) {
    hipLaunchKernelGGL(k_signs, dim3(grid_for(n)), dim3(TPB), 0, st, n, s, x, y);
}

template <typename T>
void fix_isolated(hipStream_t st, int m, const int* rows, const double* a, const T* f, T* x) {
    if (m <= 0) return;
    hipLaunchKernelGGL((k_fix_isolated<T>), dim3(grid_for(m)), dim3(TPB), 0, st, m, rows, a, f, x);
}
template void fix_isolated<double>(hipStream_t, int, const int*, const double*, const double*, double*);
template void fix_isolated<float>(hipStream_t, int, const int*, const double*, const float*, float*);

void restore_if_zero(hipStream_t st, int n, double* x, const double* scal, int slot, int m, const int* rows,
                     const double* a, const double* f) {
    hipLaunchKernelGGL(k_zero_if, dim3(grid_for(n)), dim3(TPB), 0, st, n, x, scal, slot);
    if (m > 0)
        hipLaunchKernelGGL(k_fix_isolated_if, dim3(grid_for(m)), dim3(TPB), 0, st, m, rows, a, f, x, scal, slot);
}

void elast_grad(hipStream_t st, int ne, const int* tri, const double* xy, int nn, const double* u, const double* lam,
                double nu, const double* scale, int blocked, double* grad) {
    if (ne <= 0) return;
    const double l = nu / ((1 + nu) * (1 - 2 * nu));
    const double m = 1.0 / (2 * (1 + nu));
    hipLaunchKernelGGL(k_elast_grad, dim3(grid_for(ne, 8192)), dim3(TPB), 0, st, ne, tri, xy, nn, u, lam, l, m, scale,
                       blocked, grad);
}

void seg_dot(hipStream_t st, int nchunk, const int* crow, const double* a, const double* b, double* part) {
    if (nchunk > 0) hipLaunchKernelGGL(k_seg_dot, dim3(std::min(nchunk, 65535)), dim3(TPB), 0, st, nchunk, crow, a, b, part);
}
void seg_sum(hipStream_t st, int ninst, const int* ichunk, const double* part, double* out, double* old) {
    if (ninst > 0) hipLaunchKernelGGL(k_seg_sum, dim3(std::min(ninst, 65535)), dim3(TPB), 0, st, ninst, ichunk, part, out, old);
}
void bpcg_p(hipStream_t st, int nchunk, const int* crow, const int* cinst, double* P, const double* Z, const double* s,
            const double* sold, const double* active) {
    if (nchunk > 0)
        hipLaunchKernelGGL(k_bpcg_p, dim3(std::min(nchunk, 65535)), dim3(TPB), 0, st, nchunk, crow, cinst, P, Z, s, sold,
                           active);
}
void bpcg_xr(hipStream_t st, int nchunk, const int* crow, const int* cinst, double* X, double* R, const double* P,
             const double* Q, const double* s, const double* pq, const double* active) {
    if (nchunk > 0)
        hipLaunchKernelGGL(k_bpcg_xr, dim3(std::min(nchunk, 65535)), dim3(TPB), 0, st, nchunk, crow, cinst, X, R, P, Q, s, pq,
                           active);
}
void bpcg_check(hipStream_t st, int ninst, const double* rr, const double* hb, double tol, int max_iter,
                double* active, double* iters, int count) {
    if (ninst > 0)
        hipLaunchKernelGGL(k_bpcg_check, dim3(grid_for(ninst)), dim3(TPB), 0, st, ninst, rr, hb, tol, max_iter, active,
                           iters, count);
}

void rn_scaled_hessian(hipStream_t st, int n, const int* frp, const int* fci, const double* fhv, const uint8_t* fh,
                       const double* D, const double* G, const double* J, const int* mpos, double* out,
                       double* outF, int* flag) {
    if (n > 0)
        hipLaunchKernelGGL(k_rn_scaled_hessian, dim3(std::min((n + TPB / RN_LPR - 1) / (TPB / RN_LPR), 16384)),
                           dim3(TPB), 0, st, n, frp, fci, fhv, fh, D, G, J, mpos, out, outF, flag);
}

void rn_dj(hipStream_t st, int nb, const int* idx, const double* lo, const double* up, const double* G,
           const double* X, double* D, double* J) {
    if (nb > 0) hipLaunchKernelGGL(k_rn_dj, dim3(grid_for(nb)), dim3(TPB), 0, st, nb, idx, lo, up, G, X, D, J);
}
void rn_reflect(hipStream_t st, int nb, const int* idx, const double* lo, const double* up, double* X) {
    if (nb > 0) hipLaunchKernelGGL(k_rn_reflect, dim3(grid_for(nb)), dim3(TPB), 0, st, nb, idx, lo, up, X);
}
void rn_snap(hipStream_t st, int nb, const int* idx, const double* lo, const double* up, double* X, double eps) {
    if (nb > 0) hipLaunchKernelGGL(k_rn_snap, dim3(grid_for(nb)), dim3(TPB), 0, st, nb, idx, lo, up, X, eps);
}
int rn_touch(hipStream_t st, int nb, const int* idx, const double* lo, const double* up, const double* X,
             double* part) {
    if (nb <= 0) return 0;
    const int g = std::min(grid_for(nb), 1024);
    hipLaunchKernelGGL(k_rn_touch, dim3(g), dim3(TPB), 0, st, nb, idx, lo, up, X, part);
    return g;
}
void mul(hipStream_t st, int n, double* x, const double* y) {
    hipLaunchKernelGGL(k_mul, dim3(grid_for(n)), dim3(TPB), 0, st, n, x, y);
}
void scale(hipStream_t st, int n, double* x, double a) {
    hipLaunchKernelGGL(k_scale, dim3(grid_for(n)), dim3(TPB), 0, st, n, x, a);
}

void scale_by(hipStream_t st, int n, double* x, const double* scal, int slot, int inverse) {
    hipLaunchKernelGGL(k_scale_by, dim3(grid_for(n)), dim3(TPB), 0, st, n, x, scal, slot, inverse);
}

void pcg_p(hipStream_t st, int n, double* P, const double* Z, const double* scal, int first) {
    hipLaunchKernelGGL(k_pcg_p, dim3(grid_for(n)), dim3(TPB), 0, st, n, P, Z, scal, first);
}

template <typename XT>
int pcg_xr(hipStream_t st, int n, double* X, double* R, const double* P, const double* Q, const double* scal,
           double* part, XT* ff, XT* xf, const double* sg, const uint32_t* isb, const double* isa) {
    int g = grid_for(n, 1024);
    hipLaunchKernelGGL((k_pcg_xr<XT>), dim3(g), dim3(TPB), 0, st, n, X, R, P, Q, scal, part, ff, xf, sg, isb, isa);
    return g;
}
template int pcg_xr<double>(hipStream_t, int, double*, double*, const double*, const double*, const double*, double*,
                            double*, double*, const double*, const uint32_t*, const double*);
template int pcg_xr<float>(hipStream_t, int, double*, double*, const double*, const double*, const double*, double*,
                           float*, float*, const double*, const uint32_t*, const double*);

void axpy_s(hipStream_t st, int n, double* x, const double* y, const double* scal, int slot_a, int slot_b,
            double sign) {
    hipLaunchKernelGGL(k_axpy_s, dim3(grid_for(n)), dim3(TPB), 0, st, n, x, y, scal, slot_a, slot_b, sign);
}

void axpby(hipStream_t st, int n, double a, double* x, double b, const double* y) {
    hipLaunchKernelGGL(k_axpby, dim3(grid_for(n)), dim3(TPB), 0, st, n, a, x, b, y);
}

template <typename VT, typename XT>
void mcsor_colour(hipStream_t st, const DMat<VT>& Ac, int blk0, int nblk, const int* perm, const XT* f, XT* x,
                  XT omega) {
    if (nblk <= 0) return;
    LINA_LPR_DISPATCH(Ac.lpr, hipLaunchKernelGGL((k_mcsor<VT, XT, L>), dim3(nblk), dim3(TPB), 0, st, nblk, blk0,
                                                 Ac.bk, Ac.rp, Ac.ci, Ac.cs, Ac.cb, Ac.v, Ac.dpos, perm, f, x, omega, Ac.bspec,
                                                 Ac.active))
}
template void mcsor_colour<double, double>(hipStream_t, const DMat<double>&, int, int, const int*, const double*,
                                           double*, double);
template void mcsor_colour<float, float>(hipStream_t, const DMat<float>&, int, int, const int*, const float*, float*,
                                         float);

// the first colour's sweep from a zero iterate, with the iterate's other
// entries zeroed: every product of the sweep is a_ij * 0, so row r of the
// first colour gets x_i = (1 - omega) * 0 + omega * f_i / a_ii (k_mcsor's
// value, up to the sign of a zero) and every other row 0
template <typename VT, typename XT>
__global__ void k_mcsor_first(int n, int r1, const int* __restrict__ perm, const int* __restrict__ dpos,
                              const VT* __restrict__ v, const XT* __restrict__ f, XT* __restrict__ x, XT omega) {
    for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x) {
        const int i = perm[r];
        XT xi = (XT)0;
        if (r < r1) {
            const int dp = dpos[r];
            const XT d = dp >= 0 ? (XT)v[dp] : (XT)0;
            xi = ((XT)1 - omega) * (XT)0 + omega * f[i] / d;
        }
        x[i] = xi;
    }
}

template <typename VT, typename XT>
void mcsor_first(hipStream_t st, const DMat<VT>& Ac, int rows0, const int* perm, const XT* f, XT* x, XT omega) {
    if (Ac.n == 0) return;
    hipLaunchKernelGGL((k_mcsor_first<VT, XT>), dim3(grid_for(Ac.n)), dim3(TPB), 0, st, Ac.n, rows0, perm, Ac.dpos,
                       Ac.v, f, x, omega);
}
template void mcsor_first<double, double>(hipStream_t, const DMat<double>&, int, const int*, const double*, double*,
                                          double);
template void mcsor_first<float, float>(hipStream_t, const DMat<float>&, int, const int*, const float*, float*,
                                        float);

template <typename VT, typename XT>
void sor_lex(hipStream_t st, const DMat<VT>& A, int nlv, const int* lv_off, const int* lv_rows, const XT* f, XT* x,
             XT* xo, XT omega, int sweeps, int backward) {
    if (A.n == 0) return;
    hipLaunchKernelGGL((k_sor_lex<VT, XT>), dim3(1), dim3(1024), 0, st, A.n, A.rp, A.ci, A.v, nlv, lv_off, lv_rows,
                       f, x, xo, omega, sweeps, backward);
}
template void sor_lex<double, double>(hipStream_t, const DMat<double>&, int, const int*, const int*, const double*,
                                      double*, double*, double, int, int);
template void sor_lex<float, float>(hipStream_t, const DMat<float>&, int, const int*, const int*, const float*,
                                    float*, float*, float, int, int);

template <typename VT, typename XT>
void jacobi(hipStream_t st, const DMat<VT>& A, int kind, const XT* f, const XT* xo, XT* x, int nlim) {
    if (A.nblk == 0) return;
    dim3 g(A.nblk), t(TPB);
    if (kind == JAC_REF) {
        LINA_LPR_DISPATCH(A.lpr, hipLaunchKernelGGL((k_jacobi<VT, XT, JAC_REF, L>), g, t, 0, st, A.nblk, A.bk, A.rp,
                                                    A.ci, A.cs, A.cb, A.v, A.dpos, A.l1, f, xo, x, nlim, A.bspec,
                                                    A.active))
    } else {
        LINA_LPR_DISPATCH(A.lpr, hipLaunchKernelGGL((k_jacobi<VT, XT, JAC_L1, L>), g, t, 0, st, A.nblk, A.bk, A.rp,
                                                    A.ci, A.cs, A.cb, A.v, A.dpos, A.l1, f, xo, x, nlim, A.bspec,
                                                    A.active))
    }
}
template void jacobi<double, double>(hipStream_t, const DMat<double>&, int, const double*, const double*, double*,
                                     int);

template <typename VT, typename XT>
void jacobi0(hipStream_t st, const DMat<VT>& A, int kind, const XT* f, XT* x, int nlim) {
    if (A.n == 0) return;
    const int g = std::min((A.n + TPB - 1) / TPB, 8192);
    if (kind == JAC_REF)
        hipLaunchKernelGGL((k_jacobi0<VT, XT, JAC_REF>), dim3(g), dim3(TPB), 0, st, A.n, A.dpos, A.v, A.l1, f, x,
                           nlim);
    else
        hipLaunchKernelGGL((k_jacobi0<VT, XT, JAC_L1>), dim3(g), dim3(TPB), 0, st, A.n, A.dpos, A.v, A.l1, f, x,
                           nlim);
}
template void jacobi0<double, double>(hipStream_t, const DMat<double>&, int, const double*, double*, int);
template void jacobi0<float, float>(hipStream_t, const DMat<float>&, int, const float*, float*, int);
template void jacobi<float, float>(hipStream_t, const DMat<float>&, int, const float*, const float*, float*, int);

template <typename XT>
void coarse_solve(hipStream_t st, int nd, const int* didx, const double* inv, int nt, const int* tidx,
                  const double* tdiag, const XT* b, XT* x, double* bd, bool gathered) {
    if (nd > 0 && !gathered)
        hipLaunchKernelGGL((k_coarse_gather<XT>), dim3(std::min((nd + TPB - 1) / TPB, 1024)), dim3(TPB), 0, st, nd,
                           didx, b, bd);
    int rows_blocks = (nd * WAVE + TPB - 1) / TPB;
    int tb = (nt + TPB - 1) / TPB;
    int g = rows_blocks > tb ? rows_blocks : tb;
    if (g < 1) g = 1;
    if (g > 2048) g = 2048;
    hipLaunchKernelGGL((k_coarse<XT>), dim3(g), dim3(TPB), 0, st, nd, didx, inv, nt, tidx, tdiag, b, x,
                       (const double*)bd);
}
template void coarse_solve<double>(hipStream_t, int, const int*, const double*, int, const int*, const double*,
                                   const double*, double*, double*, bool);
template void coarse_solve<float>(hipStream_t, int, const int*, const double*, int, const int*, const double*,
                                  const float*, float*, double*, bool);

template <typename XT, typename TT>
void coarse_solve_sym(hipStream_t st, int nd, int nct, int ntiles, const int* ti, const int* tj, const TT* tiles,
                      double* part, const int* didx, int nt, const int* tidx, const double* tdiag, const XT* b, XT* x,
                      double* bd, bool gathered) {
    if (nd > 0 && !gathered)
        hipLaunchKernelGGL((k_coarse_gather<XT>), dim3(std::min((nd + TPB - 1) / TPB, 1024)), dim3(TPB), 0, st, nd,
                           didx, b, bd);
    if (ntiles > 0)
        hipLaunchKernelGGL((k_csym_tiles<TT>), dim3(ntiles), dim3(TPB), 0, st, nct, ti, tj, tiles, bd, part);
    const int g = std::max(std::max(nct, (nt + TPB - 1) / TPB), 1);
    hipLaunchKernelGGL((k_csym_reduce<XT>), dim3(g), dim3(TPB), 0, st, nd, nct, part, didx, nt, tidx, tdiag, b, x);
}
template void coarse_solve_sym<double, double>(hipStream_t, int, int, int, const int*, const int*, const double*,
                                               double*, const int*, int, const int*, const double*, const double*,
                                               double*, double*, bool);
template void coarse_solve_sym<float, double>(hipStream_t, int, int, int, const int*, const int*, const double*,
                                              double*, const int*, int, const int*, const double*, const float*,
                                              float*, double*, bool);

template <typename TT>
void inverse_to_tiles(hipStream_t st, int nd, int ntiles, const int* ti, const int* tj, const double* inv,
                      TT* tiles) {
    if (ntiles <= 0) return;
    hipLaunchKernelGGL((k_inv_tiles<TT>), dim3(4096), dim3(TPB), 0, st, nd, ntiles, ti, tj, inv, tiles);
}
template void inverse_to_tiles<double>(hipStream_t, int, int, const int*, const int*, const double*, double*);

void trsv(hipStream_t st, const int* rp, const int* ci, const double* v, int nlv, const int* lv_off,
          const int* lv_rows, const double* z, double* w, int upper) {
    hipLaunchKernelGGL(k_trsv, dim3(1), dim3(1024), 0, st, rp, ci, v, nlv, lv_off, lv_rows, z, w, upper);
}

void tri_solve_pair(hipStream_t st, const TriPack& fw, const TriPack& bw, const int* perm, const double* z,
                    double* zp, double* y, double* w) {
    if (fw.n <= 0) return;
    hipLaunchKernelGGL(k_gather_perm, dim3(std::min(1024, (fw.n + 255) / 256)), dim3(256), 0, st, fw.n, perm, z, zp);
    const int maxp = std::max(fw.maxp, bw.maxp);
    if (maxp > 4)
        hipLaunchKernelGGL((k_tri_pair<8, 2>), dim3(1), dim3(TRI_NT), 0, st, fw, bw, zp, y, w);
    else if (maxp == 4)
        hipLaunchKernelGGL((k_tri_pair<4, 2>), dim3(1), dim3(TRI_NT), 0, st, fw, bw, zp, y, w);
    else
        hipLaunchKernelGGL((k_tri_pair<3, 2>), dim3(1), dim3(TRI_NT), 0, st, fw, bw, zp, y, w);
}

void sor_packed(hipStream_t st, const SorPack& s, const double* f, double* x, double omega, int sweeps) {
    const int n = s.p.n;
    if (n <= 0) return;
    const SorArgs so{s.a, s.prod, omega, s.ocnt, s.ooff, s.prodov, s.old_first};
    for (int k = 0; k < sweeps; k++) {
        hipLaunchKernelGGL(k_sor_gather, dim3(std::min(1024, (n + 255) / 256)), dim3(256), 0, st, n, s.p.maxp, s.row,
                           s.ocnt, s.oval, s.ocol, s.ooff, s.oovv, s.oovc, s.old_first, f, (const double*)x, omega,
                           s.in, s.a, s.prod, s.prodov);
        if (s.p.maxp > 4)
            hipLaunchKernelGGL((k_sor_sweep<8, 2>), dim3(1), dim3(TRI_NT), 0, st, s.p, so, (const double*)s.in, x);
        else
            hipLaunchKernelGGL((k_sor_sweep<4, 2>), dim3(1), dim3(TRI_NT), 0, st, s.p, so, (const double*)s.in, x);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) throw std::runtime_error(std::string("k_sor_sweep launch: ") + hipGetErrorString(e));
    }
}

void tri_pack_values(hipStream_t st, int n, int maxp, const int* vmap, const int* dmap, int nov, const int* ovmap,
                     const double* v, void* rec, double* ovval) {
    const int m = std::max(n, nov);
    if (m <= 0) return;
    const dim3 g(std::min(4096, (m + 255) / 256)), b(256);
    if (maxp == 3)
        hipLaunchKernelGGL(k_tri_values<3>, g, b, 0, st, n, vmap, dmap, nov, ovmap, v, (TriRec<3>*)rec, ovval);
    else if (maxp <= 4)
        hipLaunchKernelGGL(k_tri_values<4>, g, b, 0, st, n, vmap, dmap, nov, ovmap, v, (TriRec<4>*)rec, ovval);
    else
        hipLaunchKernelGGL(k_tri_values<8>, g, b, 0, st, n, vmap, dmap, nov, ovmap, v, (TriRec<8>*)rec, ovval);
}

void ilu0_planned(hipStream_t st, int nlv, const int* lv_off, const int4* m1, const int4* m2, const int* mask,
                  const int* pu, const double* vA, double* vL, double* vU, int* fail) {
    hipLaunchKernelGGL(k_ilu0_planned, dim3(1), dim3(1024), 0, st, nlv, lv_off, m1, m2, mask, pu, vA, vL, vU, fail);
}

void ilu0_planned_levels(hipStream_t st, const std::vector<int>& off, const int4* m1, const int4* m2,
                         const int* mask, const int* pu, const double* vA, double* vL, double* vU, int* fail) {
    for (size_t L = 0; L + 1 < off.size(); L++) {
        const int cnt = off[L + 1] - off[L];
        if (cnt <= 0) continue;
        hipLaunchKernelGGL(k_ilu0_planned_level, dim3(grid_for(cnt)), dim3(TPB), 0, st, off[L], off[L + 1], m1, m2,
                           mask, pu, vA, vL, vU, fail);
    }
}

void tri_levels_wide(hipStream_t st, const int* rp, const int* ci, const double* v, const std::vector<int>& off,
                     const int* rows, const double* in, double* out, int upper) {
    for (size_t L = 0; L + 1 < off.size(); L++) {
        const int cnt = off[L + 1] - off[L];
        if (cnt <= 0) continue;
        hipLaunchKernelGGL(k_tri_level, dim3(grid_for(cnt)), dim3(TPB), 0, st, off[L], off[L + 1], rows, rp, ci, v,
                           in, out, upper);
    }
}

void permute_idx(hipStream_t st, int n, const int* idx, const double* in, double* out) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_permute, dim3(grid_for(n)), dim3(TPB), 0, st, n, idx, in, out, 0);
}

void permute(hipStream_t st, int n, const int* perm, const double* in, double* out, int inverse) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_permute, dim3(grid_for(n)), dim3(TPB), 0, st, n, perm, in, out, inverse);
}

void ilu0_numeric(hipStream_t st, int nlv, const int* lv_off, const int* lv_rows, const int* rpA, const int* ciA,
                  const double* vA, const int* rpL, const int* ciL, double* vL, const int* rpU, const int* ciU,
                  double* vU, int* fail) {
    hipLaunchKernelGGL(k_ilu0, dim3(1), dim3(1024), 0, st, nlv, lv_off, lv_rows, rpA, ciA, vA, rpL, ciL, vL, rpU, ciU,
                       vU, fail);
}

int multidot(hipStream_t st, int n, int m, const double* V, size_t ld, const double* r, double* part, double* out) {
    int g = grid_for(n, 512);
    hipLaunchKernelGGL(k_multidot, dim3(g), dim3(TPB), 0, st, n, m, V, ld, r, part);
    hipLaunchKernelGGL(k_multidot_final, dim3(1), dim3(TPB), 0, st, m, g, part, out);
    return g;
}

void combine(hipStream_t st, int n, int m, const double* V, size_t ld, const double* c, const double* base,
             double sign, double* y) {
    hipLaunchKernelGGL(k_combine, dim3(grid_for(n)), dim3(TPB), 0, st, n, m, V, ld, c, base, sign, y);
}

// piv: one device double of scratch
template <typename XT>
void bcoarse_solve(hipStream_t st, int nd, const int* rblk, const int* off, const size_t* moff, const int* didx,
                   const double* inv, int nt, const int* tidx, const double* tdiag, const XT* b, XT* x,
                   const double* active) {
    const int waves_per_block = TPB / WAVE;
    int g = std::max(1, std::min((nd + waves_per_block - 1) / waves_per_block, 8192));
    g = std::max(g, std::min((nt + TPB - 1) / TPB, 4096));
    hipLaunchKernelGGL((k_bcoarse<XT>), dim3(g), dim3(TPB), 0, st, nd, rblk, off, moff, didx, inv, nt, tidx, tdiag,
                       b, x, active);
}
template void bcoarse_solve<double>(hipStream_t, int, const int*, const int*, const size_t*, const int*,
                                    const double*, int, const int*, const double*, const double*, double*,
                                    const double*);
template void bcoarse_solve<float>(hipStream_t, int, const int*, const int*, const size_t*, const int*,
                                   const double*, int, const int*, const double*, const float*, float*,
                                   const double*);

void batched_gauss_jordan_inverse(hipStream_t st, int nd, int max_block, const int* rblk, const int* off,
                                  const size_t* moff, double* M, double* piv) {
    if (nd <= 0) return;
    const int gw = std::max(1, std::min((nd + TPB / WAVE - 1) / (TPB / WAVE), 8192));
    const int gp = std::max(1, std::min((nd + TPB - 1) / TPB, 4096));
    for (int k = 0; k < max_block; k++) {
        hipLaunchKernelGGL(k_bgj_update, dim3(gw), dim3(TPB), 0, st, nd, k, rblk, off, moff, M, piv);
        hipLaunchKernelGGL(k_bgj_pivot, dim3(gp), dim3(TPB), 0, st, nd, k, rblk, off, moff, M, (const double*)piv);
    }
}

size_t gauss_jordan_work(int N) { return (size_t)GJB * GJB + 2 * (size_t)GJB * (size_t)std::max(N, 1); }

void gauss_jordan_inverse(hipStream_t st, int N, double* M, double* work) {
    if (N <= 0) return;
    const char* e = getenv("LINA_GJ");  // read per inversion (rare)
    if (e && std::strcmp(e, "scalar") == 0) {  // the unblocked steps, for comparisons
        int g = std::min(N, 4096);
        int gp = (N + TPB - 1) / TPB;
        for (int k = 0; k < N; k++) {
            hipLaunchKernelGGL(k_gj_update, dim3(g), dim3(TPB), 0, st, N, k, M, work);
            hipLaunchKernelGGL(k_gj_pivot, dim3(gp), dim3(TPB), 0, st, N, k, M, (const double*)work);
        }
        return;
    }
    double* Dinv = work;
    double* Acol = work + GJB * GJB;
    double* Brow = Acol + (size_t)GJB * N;
    const int nt = (N + GJB - 1) / GJB;
    for (int kb = 0; kb < nt; kb++) {
        const int k0 = kb * GJB, nk = std::min(GJB, N - k0);
        hipLaunchKernelGGL(k_bgj_diag, dim3(1), dim3(1024), 0, st, N, k0, nk, M, Dinv);
        hipLaunchKernelGGL(k_bgj_panels, dim3(2 * nt), dim3(256), 0, st, N, kb, nt, nk, M, (const double*)Dinv, Acol,
                           Brow);
        if (nt > 1)
            hipLaunchKernelGGL(k_bgj_gemm, dim3((nt - 1) * (nt - 1)), dim3(256), 0, st, N, kb, nt, M,
                               (const double*)Acol, (const double*)Brow);
    }
}

void plan_sum(hipStream_t st, int ne, const int64_t* ptr, const int* ia, const int* ib, const uint8_t* seed,
              const double* va, const double* vb, double* out) {
    if (ne > 0)
        hipLaunchKernelGGL(k_plan_sum_staged, dim3((ne + TPB - 1) / TPB), dim3(TPB), 0, st, ne, ptr, ia, ib, seed, va,
                           vb, out);
}

void plan_scatter(hipStream_t st, int ne, const double* val, const int* pos, const uint8_t* lone, double* dst,
                  int* flag) {
    if (ne > 0) hipLaunchKernelGGL(k_plan_scatter, dim3(grid_for(ne)), dim3(TPB), 0, st, ne, val, pos, lone, dst, flag);
}

void dirichlet_rows(hipStream_t st, int n, const int* rp, const int* ci, const uint8_t* dir, const double* u,
                    double* val, double* b) {
    if (n > 0) hipLaunchKernelGGL(k_dirichlet_rows, dim3(grid_for(n)), dim3(TPB), 0, st, n, rp, ci, dir, u, val, b);
}

void negate_rows(hipStream_t st, int m, const int* rows, const int* rp, double* v) {
    if (m > 0) hipLaunchKernelGGL(k_negate_rows, dim3(std::min(m, 4096)), dim3(64), 0, st, m, rows, rp, v);
}

template <typename VT>
void row_l1(hipStream_t st, int n, const int* rp, const double* v, VT* out) {
    if (n > 0) hipLaunchKernelGGL((k_row_l1<VT>), dim3(grid_for(n)), dim3(TPB), 0, st, n, rp, v, out);
}
template void row_l1<double>(hipStream_t, int, const int*, const double*, double*);
template void row_l1<float>(hipStream_t, int, const int*, const double*, float*);

template <typename T>
void gather(hipStream_t st, int n, const int* idx, const T* src, T* dst) {
    if (n > 0) hipLaunchKernelGGL((k_gather<T>), dim3(grid_for(n)), dim3(TPB), 0, st, n, idx, src, dst);
}
template void gather<double>(hipStream_t, int, const int*, const double*, double*);
template void gather<float>(hipStream_t, int, const int*, const float*, float*);

void scatter(hipStream_t st, int n, const int* idx, const double* src, double* dst) {
    if (n > 0) hipLaunchKernelGGL(k_scatter, dim3(grid_for(n)), dim3(TPB), 0, st, n, idx, src, dst);
}

}  // namespace dev
}  // namespace lina
