// Plans of the device Galerkin products (rap.cpp).
#pragma once

#include <cstdint>
#include <vector>

#include "lina_internal.hpp"

namespace lina {

// The products of C = A B in LinA's traversal, grouped by entry of C.  S
// holds C's structural pattern (rows sorted by column; values unused); entry
// e receives the products A.v[ia[t]] * B.v[ib[t]] for t in [ptr[e],
// ptr[e+1]) in that order, starting from +0.0 when seed[e] (the row's
// sparse_zero diagonal), else from its first product.
struct ProductPlan {
    Csr S;
    std::vector<int64_t> ptr;
    std::vector<uint8_t> seed;
    std::vector<int> ia, ib;
};

// a_all / b_all: list every structural entry of that factor (its values
// change between uses of the plan); otherwise only its nonzero entries (its
// values are fixed).  Entries of C in columns >= ncols are not formed.
void plan_product(const Csr& A, bool a_all, const Csr& B, bool b_all, int ncols, ProductPlan& P);

// R (A P) of one level: ap = A P, ac = R ap; pos[e]: position of ac's entry
// e in the coarse level's CSR (-1: not in its pattern); lone[e]: that row of
// the coarse CSR is the single entry (i, 0.0) delete_zeros leaves in a row
// with nothing else
struct RapLevelPlan {
    ProductPlan ap, ac;
    std::vector<int> pos;
    std::vector<uint8_t> lone;
};
void plan_galerkin(const HostLevel& fine, const Csr& Ac, RapLevelPlan& L);
size_t plan_bytes(const ProductPlan& P);

}  // namespace lina
