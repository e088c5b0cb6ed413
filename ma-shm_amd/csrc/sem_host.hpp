// Host-side building blocks shared by the library's translation units:
// GLL rules, the Lagrange derivative matrix and the node->(element,point) maps.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace semhost {

// Occurrence records of the patch plans (PatchPlan::occ8): entries per node, the end mark
// of a record with fewer occurrences, and the mark of a node whose further occurrences are
// in its occurrence list.  Occurrence values (LDS point indices) stay below kOccMore.
constexpr int kOccRec = 8;
constexpr uint16_t kOccEnd = 0xFFFF;
constexpr uint16_t kOccMore = 0xFFFE;

// GLL nodes and weights for n points on [-1,1] (gll.m:1-39): the reference's literal table
// for n = 2..10; beyond it the closed form (interior nodes the roots of P'_{n-1},
// w = 2/(n(n-1) P_{n-1}(x)^2)).
void gll(int n, std::vector<double>& x, std::vector<double>& w);

// D[i*n+j] = L_j'(x_i): derivative of the j-th Lagrange basis polynomial at node i.
// Equals Sxprim of shape2D_prim.m:41-49 (which builds it through inv(Vandermonde)).
void lagrange_derivative(int n, const std::vector<double>& x, std::vector<double>& D);

// 12-basket colouring (nodesMaps_Fiborek.m:1-27).  For every element-local point
// q = e*n_loc + i returns basket[q] in [0,12) and row[q] in that basket.
// Fails (returns false) if a node occurs more than 12 times.
// Even-odd form of an n x n centro-antisymmetric matrix M (M[n-1-i][n-1-m] = -M[i][m], as the
// GLL derivative matrix and its transpose are), appended to out: with h = n / 2, A[i][m] =
// (M[i][m] - M[i][n-1-m]) / 2 and B[i][m] = (M[i][m] + M[i][n-1-m]) / 2 for i, m < h; for odd n
// also M[i][h] (i < h) and M[h][m] (m < h).  Then, with e[m] = u[m] + u[n-1-m] and
// o[m] = u[m] - u[n-1-m], (M u)[i] = P + Q and (M u)[n-1-i] = P - Q for P = sum A[i][m] o[m],
// Q = sum B[i][m] e[m] (+ M[i][h] u[h]), and (M u)[h] = sum M[h][m] o[m].
void even_odd_table(int n, const std::vector<double>& M, std::vector<double>& out);

bool basket_colouring(int64_t n_el, int n_loc, const int32_t* nodes, int64_t n_nodes,
                      std::vector<int8_t>& basket, std::vector<int32_t>& row,
                      int64_t& rows);

// Node -> occurrences (q = e*n_loc + i) in CSR, ordered by basket index, which is
// the order in which the reference's basket loop adds the contributions
// (force_excit_c.m:666-668).
void inverse_map(int64_t n_el, int n_loc, const int32_t* nodes, int64_t n_nodes,
                 const std::vector<int8_t>& basket, std::vector<int32_t>& rowptr,
                 std::vector<int32_t>& occ);

// Element patches for the fused element/assembly kernel.  Elements are ordered by the
// Morton code of their (x, y) centroid cell with the through-thickness layer as minor key,
// and cut into chunks of P: a chunk of 16 shells is a 4x4 block, a chunk of 8 two-layer
// hexahedra a 2x2x2 block.  Per patch: its nodes (owned ones -- every occurrence inside the
// patch -- first), the occurrences of each node in basket order as LDS point indices, and
// one partial-sum slot per boundary node (slots patch-major: slot_base[b] + position among
// the patch's boundary nodes); per boundary node: its slots in patch order.
// A patch's local map, occurrence ranges and occurrence lists depend only on its local
// structure, and the interior patches of a regular mesh share them: they are stored once per
// distinct structure (a template) and each patch records the offsets of its template's
// three blocks (lm_base, optr_base, occ_base).
// Nodes get an internal numbering in patch order (each patch's nodes in its local order,
// owned first, numbered when first met: a patch's owned nodes are consecutive), so that a
// patch reads and writes mostly contiguous node records; pn_id and sh_id hold internal ids.
struct PatchPlan {
  int P = 0;
  int64_t n_patch = 0;
  std::vector<int32_t> perm;      // new element -> caller's element
  std::vector<int32_t> el0;       // [n_patch+1]
  std::vector<int32_t> pn_base;   // [n_patch+1]
  std::vector<int32_t> pn_owned;  // [n_patch]
  std::vector<int32_t> pn_id;
  std::vector<int32_t> slot_base; // [n_patch]
  std::vector<uint16_t> occ_ptr;  // per template pn+1 occurrence offsets
  std::vector<uint16_t> occ;      // per template its occurrence lists (LDS point le*n_loc + p)
  std::vector<uint16_t> lmap;     // per template [elements][n_loc] patch-local node of each point
  std::vector<int32_t> optr_base; // [n_patch] offset of the patch's template in occ_ptr
  std::vector<int32_t> occ_base;  // [n_patch] ... in occ
  std::vector<int32_t> lm_base;   // [n_patch] ... in lmap
  // per template [pn][kOccRec] occurrence records (the kernels read a node's record with one
  // 16-byte load): a node's occurrences in list order, unused entries kOccEnd; a node with
  // more than kOccRec occurrences holds its first kOccRec - 1 and then kOccMore (the rest
  // from its occurrence list)
  std::vector<uint16_t> occ8;
  std::vector<int32_t> o8_base;   // [n_patch] ... in occ8
  int64_t n_templates = 0;
  std::vector<int32_t> sh_id, sh_ptr, sh_slot;
  std::vector<int32_t> node_new;  // caller's node -> internal node
  std::vector<int32_t> node_old;  // internal node -> caller's node
  int max_pn = 0;
  int64_t n_slots = 0;
};

// group (nullable): per element a case id, the major ordering key, so that no patch mixes
// cases of a batch
// Uniform-grid sampling of a flat-shell mesh (spec2meshgrid_flat_shell.m:103-299): grid
// point j = r + Ny*c at (xi[c], yi[r]); its owner element among the first n_search
// elements (bounding-box candidates in element order; the only one, else the first
// strictly inside, else the first); its local coordinates (s, t) by the reference's
// Newton iteration (a step with the inverse Jacobian at the element's first node from the
// nearest element node, then three steps at the current estimate).  Lx/Ly hold the 1D
// Lagrange weights L_i(s) and L_j(t) of every point ([P][n] each).
struct GridPlan {
  int Nx = 0, Ny = 0, n = 0;
  std::vector<double> xi, yi;
  std::vector<int32_t> owner;
  std::vector<double> s, t, Lx, Ly;
};
bool grid_plan(int64_t n_el, int n, const int32_t* nodes, int64_t n_nodes, const double* coords,
               int64_t n_search, int Nx, int Ny, GridPlan& g, std::string& err);

bool build_patches(int64_t n_el, int n_loc, const int32_t* nodes, int64_t n_nodes,
                   const double* coords, int P, const std::vector<int8_t>& basket,
                   PatchPlan& plan, const int32_t* group = nullptr);

}  // namespace semhost
