// Re-assembly of the degraded stiffness on the device.
//
// Crack's elastic Hessian is the contraction of the third-order tensor
// Strain_grad (B, rows = nodes k, triples (n b + j, n d + k', p)) with the
// nodal rank-4 moduli C (component (a,b,c,d) of node k at n (8a+4b+2c+d) + k):
// sparse3D_times_rank4_2D_right_use_map (FEM2D/FEM2D.c:8534-8600) adds, for
// every node k, every triple of its row in order and (a, c) in the order
// (0,0), (0,1), (1,1), (1,0), the product C[n (8a+4b+2c+d) + k] * p to the
// entry (n a + j, n c + k') of a matrix that starts at zero.  Its map
// (sparse3D_times_rank4_2D_right_get_map, :8504-8532) fixes the target
// positions once per pattern.
//
// Here the plan lists, per entry of the target CSR, its contributions in
// that order (an entry's sum is then bit-identical to the reference's
// sequential accumulation), built once on the host; every assembly is one
// device gather-sum over the plan (dev::plan_sum), optionally followed by
// the symmetric Dirichlet treatment of the rows (dev::dirichlet_rows), so a
// new damage field becomes new stiffness values without leaving the GPU.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <stdexcept>
#include <vector>

#include "../../include/lina_hip.h"
#include "kernels.hpp"
#include "lina_internal.hpp"
#include "solver.hpp"

struct lina_assembly {
    int n = 0;     // nodes
    int N = 0;     // rows of the target (2 n)
    int nnz = 0;   // target entries
    int64_t nq = 0;
    int *rp = nullptr, *ci = nullptr;
    int64_t* ptr = nullptr;
    int *ic = nullptr, *ib = nullptr;
    uint8_t* seed = nullptr;
    double* bv = nullptr;
    uint8_t* dir = nullptr;
    hipStream_t st = nullptr;
    lina::DevBuffers bufs;
    ~lina_assembly() {
        if (st) (void)hipStreamDestroy(st);
    }
};

namespace {

template <typename F>
int aguard(F&& f) {
    try {
        return f();
    } catch (const lina::HipError& e) {
        lina::set_error("%s", e.what());
        return LINA_EHIP;
    } catch (const std::invalid_argument& e) {
        lina::set_error("%s", e.what());
        return LINA_EARG;
    } catch (const std::exception& e) {
        lina::set_error("%s", e.what());
        return LINA_EARG;
    }
}

}  // namespace

extern "C" {

lina_assembly* lina_assembly_create(int n, const int* bptr, const int* bi1, const int* bi2, const double* bval,
                                    const int* rowptr, const int* col) {
    lina_assembly* h = nullptr;
    int rc = aguard([&]() {
        if (n <= 0 || !bptr || !rowptr || !col) throw std::invalid_argument("lina_assembly_create: bad arguments");
        const int N = 2 * n;
        const int64_t nb = bptr[n];
        if (bptr[0] != 0 || nb < 0 || (nb > 0 && (!bi1 || !bi2 || !bval)))
            throw std::invalid_argument("lina_assembly_create: bad tensor row offsets");
        for (int k = 0; k < n; k++)
            if (bptr[k + 1] < bptr[k]) throw std::invalid_argument("lina_assembly_create: tensor rows out of order");
        if (rowptr[0] != 0) throw std::invalid_argument("lina_assembly_create: rowptr[0] != 0");
        for (int r = 0; r < N; r++) {
            if (rowptr[r + 1] < rowptr[r]) throw std::invalid_argument("lina_assembly_create: rowptr decreases");
            for (int p = rowptr[r]; p < rowptr[r + 1]; p++)
                if (col[p] < 0 || col[p] >= N || (p > rowptr[r] && col[p] <= col[p - 1]))
                    throw std::invalid_argument("lina_assembly_create: columns out of range or not sorted");
        }
        const int nnz = rowptr[N];
        auto position = [&](int r, int c) {
            const int* b = col + rowptr[r];
            const int* e = col + rowptr[r + 1];
            const int* it = std::lower_bound(b, e, c);
            if (it == e || *it != c)
                throw std::invalid_argument("lina_assembly_create: an entry of the contraction is not in the pattern");
            return (int)(it - col);
        };
        // contributions in the reference's order: target position, C index, B index
        const int ac[4][2] = {{0, 0}, {0, 1}, {1, 1}, {1, 0}};
        std::vector<int> tgt((size_t)nb * 4), cidx((size_t)nb * 4), bidx((size_t)nb * 4);
        std::vector<int64_t> cnt((size_t)nnz + 1, 0);
        size_t t = 0;
        for (int k = 0; k < n; k++)
            for (int q = bptr[k]; q < bptr[k + 1]; q++) {
                if (bi1[q] < 0 || bi1[q] >= N || bi2[q] < 0 || bi2[q] >= N)
                    throw std::invalid_argument("lina_assembly_create: tensor index beyond 2 n");
                const int b = bi1[q] < n ? 0 : 1, d = bi2[q] < n ? 0 : 1;
                const int j = bi1[q] % n, kk = bi2[q] % n;
                for (const auto& s : ac) {
                    const int a = s[0], c = s[1];
                    tgt[t] = position(n * a + j, n * c + kk);
                    cidx[t] = n * (8 * a + 4 * b + 2 * c + d) + k;
                    bidx[t] = q;
                    cnt[(size_t)tgt[t] + 1]++;
                    t++;
                }
            }
        for (int e = 0; e < nnz; e++) cnt[e + 1] += cnt[e];
        std::vector<int> ic(t), ib(t);
        std::vector<int64_t> fill(cnt.begin(), cnt.end() - 1);
        for (size_t u = 0; u < t; u++) {  // stable: each entry keeps the reference's order
            const int64_t w = fill[tgt[u]]++;
            ic[w] = cidx[u];
            ib[w] = bidx[u];
        }
        h = new lina_assembly();
        h->n = n;
        h->N = N;
        h->nnz = nnz;
        h->nq = (int64_t)t;
        lina::hip_check(hipStreamCreateWithFlags(&h->st, hipStreamNonBlocking), "hipStreamCreate");
        std::vector<int> rp(rowptr, rowptr + N + 1), ci(col, col + nnz);
        std::vector<double> bv(bval, bval + nb);
        std::vector<uint8_t> seed((size_t)nnz, 1);  // every entry starts at +0.0, as the reference's zero matrix
        h->rp = lina::dupload(rp, h->bufs, h->st);
        h->ci = lina::dupload(ci, h->bufs, h->st);
        h->ptr = lina::dupload(cnt, h->bufs, h->st);
        h->ic = lina::dupload(ic, h->bufs, h->st);
        h->ib = lina::dupload(ib, h->bufs, h->st);
        h->seed = lina::dupload(seed, h->bufs, h->st);
        h->bv = lina::dupload(bv, h->bufs, h->st);
        lina::hip_check(hipStreamSynchronize(h->st), "sync");
        return 0;
    });
    if (rc < 0) {
        delete h;
        return nullptr;
    }
    return h;
}

void lina_assembly_destroy(lina_assembly* h) { delete h; }

int lina_assembly_nnz(const lina_assembly* h) { return h ? h->nnz : LINA_EARG; }

long lina_assembly_terms(const lina_assembly* h) { return h ? (long)h->nq : (long)LINA_EARG; }

int lina_assembly_set_dirichlet(lina_assembly* h, const unsigned char* is_dirichlet) {
    return aguard([&]() {
        if (!h) throw std::invalid_argument("null handle");
        if (!is_dirichlet) {
            h->dir = nullptr;
            return 0;
        }
        std::vector<uint8_t> d(is_dirichlet, is_dirichlet + h->N);
        for (uint8_t& v : d) v = v ? 1 : 0;
        h->dir = lina::dupload(d, h->bufs, h->st);
        lina::hip_check(hipStreamSynchronize(h->st), "sync");
        return 0;
    });
}

int lina_assembly_values(lina_assembly* h, const double* C, double* val, const double* u_dir, double* b) {
    return aguard([&]() {
        if (!h || !C || !val) throw std::invalid_argument("lina_assembly_values: null argument");
        if (h->dir && !u_dir) throw std::invalid_argument("lina_assembly_values: Dirichlet values missing");
        lina::dev::plan_sum(h->st, h->nnz, h->ptr, h->ic, h->ib, h->seed, C, h->bv, val);
        if (h->dir) lina::dev::dirichlet_rows(h->st, h->N, h->rp, h->ci, h->dir, u_dir, val, b);
        lina::hip_check(hipGetLastError(), "launch");
        lina::hip_check(hipStreamSynchronize(h->st), "sync");
        return 0;
    });
}

int lina_dirichlet_symmetric(int n, const int* rowptr, const int* col, const unsigned char* is_dirichlet,
                             const double* u, double* val, double* b) {
    return aguard([&]() {
        if (n < 0 || !rowptr || !col || !is_dirichlet || !u || !val)
            throw std::invalid_argument("lina_dirichlet_symmetric: null argument");
        lina::dev::dirichlet_rows(nullptr, n, rowptr, col, (const uint8_t*)is_dirichlet, u, val, b);
        lina::hip_check(hipGetLastError(), "launch");
        lina::hip_check(hipStreamSynchronize(nullptr), "sync");
        return 0;
    });
}

}  // extern "C"
