// Host-side building blocks shared by the library's translation units:
// GLL rules, the Lagrange derivative matrix and the node->(element,point) maps.
#pragma once
#include <cstdint>
#include <vector>

namespace semhost {

// GLL nodes and weights for n points on [-1,1] (gll.m:1-39), from the closed form:
// interior nodes are the roots of P'_{n-1}, w = 2/(n(n-1) P_{n-1}(x)^2).
void gll(int n, std::vector<double>& x, std::vector<double>& w);

// D[i*n+j] = L_j'(x_i): derivative of the j-th Lagrange basis polynomial at node i.
// Equals Sxprim of shape2D_prim.m:41-49 (which builds it through inv(Vandermonde)).
void lagrange_derivative(int n, const std::vector<double>& x, std::vector<double>& D);

// 12-basket colouring (nodesMaps_Fiborek.m:1-27).  For every element-local point
// q = e*n_loc + i returns basket[q] in [0,12) and row[q] in that basket.
// Fails (returns false) if a node occurs more than 12 times.
bool basket_colouring(int64_t n_el, int n_loc, const int32_t* nodes, int64_t n_nodes,
                      std::vector<int8_t>& basket, std::vector<int32_t>& row,
                      int64_t& rows);

// Node -> occurrences (q = e*n_loc + i) in CSR, ordered by basket index, which is
// the order in which the reference's basket loop adds the contributions
// (force_excit_c.m:666-668).
void inverse_map(int64_t n_el, int n_loc, const int32_t* nodes, int64_t n_nodes,
                 const std::vector<int8_t>& basket, std::vector<int32_t>& rowptr,
                 std::vector<int32_t>& occ);

}  // namespace semhost
