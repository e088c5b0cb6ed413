// Host-side part of the library: GLL rules, index maps, mesh front end and
// laminate helpers.  Compiled with -ffp-contract=off so that the mesh front end
// evaluates its coordinate expressions exactly as written (no fused multiply-add),
// which the edge-node numbering (a sort of coordinates) depends on.
#include "sem_host.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <numeric>
#include <iterator>
#include <tuple>
#include <unordered_map>

#include "../../include/sem_hip.h"

namespace semhost {

static void legendre(int N, double x, double& P, double& dP) {
  // P_N(x) and P_N'(x) by the three-term recurrence.
  double p0 = 1.0, p1 = x, d0 = 0.0, d1 = 1.0;
  if (N == 0) { P = 1.0; dP = 0.0; return; }
  for (int k = 2; k <= N; ++k) {
    double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
    double d2 = d0 + (2.0 * k - 1.0) * p1;
    p0 = p1; p1 = p2; d0 = d1; d1 = d2;
  }
  P = p1; dP = d1;
}

// The reference's GLL table (gll.m:7-36), n = 2..10: the first ceil(n/2) nodes and weights
// as it writes them; the rest follow by symmetry (ksi(n+1-i) = -ksi(i), wi(n+1-i) = wi(i)).
// Kept to the last bit: node coordinates of the mesh front end, and with them the order
// unique(...,'rows') numbers edge nodes in, depend on these values.
struct GllRow {
  double x[5], w[5];
};
static const GllRow kGllTable[9] = {
    {{-1.0}, {1.0}},  // n = 2
    {{-1.0, 0.0}, {0.3333333333333333, 1.3333333333333333}},  // n = 3
    {{-1.0, -0.4472135954999579}, {0.16666666666666666, 0.8333333333333333}},  // n = 4
    {{-1.0, -0.6546536707079771, 0.0}, {0.1, 0.5444444444444445, 0.7111111111111111}},  // n = 5
    {{-1.0, -0.7650553239294647, -0.28523151648064504},
     {0.06666666666666667, 0.37847495629784766, 0.5548583770354865}},  // n = 6
    {{-1.0, -0.8302238962785669, -0.4688487934707142, 0.0},
     {0.047619047619047616, 0.2768260473615661, 0.43174538120986283, 0.4876190476190476}},  // n = 7
    {{-1.0, -0.8717401485096066, -0.5917001814331423, -0.2092992179024788},
     {0.03571428571428571, 0.2107042271435067, 0.3411226924835044, 0.4124587946587037}},  // n = 8
    {{-1.0, -0.8997579954114602, -0.6771862795107377, -0.36311746382617816, 0.0},
     {0.027777777777777776, 0.16549536156080175, 0.27453871250016254, 0.34642851097304594,
      0.37151927437641724}},  // n = 9
    {{-1.0, -0.9195339081664589, -0.738773865105505, -0.4779249498104445, -0.16527895766638692},
     {0.022222222222222223, 0.1333059908510802, 0.22488934206312652, 0.29204268367968383,
      0.32753976118389744}},  // n = 10
};

void gll(int n, std::vector<double>& x, std::vector<double>& w) {
  const int N = n - 1;
  x.assign(n, 0.0);
  w.assign(n, 0.0);
  if (n >= 2 && n <= 10) {
    const GllRow& t = kGllTable[n - 2];
    for (int i = 0; i < (n + 1) / 2; ++i) {
      x[i] = t.x[i];
      w[i] = t.w[i];
      if (i < n / 2) {  // the middle node of odd n stays +0
        x[N - i] = -t.x[i];
        w[N - i] = t.w[i];
      }
    }
    return;
  }
  x[0] = -1.0; x[N] = 1.0;
  for (int i = 1; i < N; ++i) {
    double xi = -std::cos(M_PI * i / N);
    for (int it = 0; it < 100; ++it) {
      double P, dP;
      legendre(N, xi, P, dP);
      // P'' from Legendre's equation (1-x^2)P'' = 2x P' - N(N+1) P
      double d2P = (2.0 * xi * dP - N * (N + 1.0) * P) / (1.0 - xi * xi);
      double dx = dP / d2P;
      xi -= dx;
      if (std::fabs(dx) < 1e-17) break;
    }
    x[i] = xi;
  }
  for (int i = 0; i < n; ++i) {  // exact antisymmetry, as the reference table
    double a = 0.5 * (x[i] - x[N - i]);
    x[i] = a;
  }
  if (n % 2) x[n / 2] = 0.0;
  x[0] = -1.0; x[N] = 1.0;
  for (int i = 0; i < n; ++i) {
    double P, dP;
    legendre(N, x[i], P, dP);
    w[i] = 2.0 / (N * (N + 1.0) * P * P);
  }
  for (int i = 0; i < n / 2; ++i) {
    double a = 0.5 * (w[i] + w[N - i]);
    w[i] = a; w[N - i] = a;
  }
}

void lagrange_derivative(int n, const std::vector<double>& x, std::vector<double>& D) {
  const int N = n - 1;
  std::vector<double> P(n);
  for (int i = 0; i < n; ++i) { double dP; legendre(N, x[i], P[i], dP); }
  D.assign((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j)
      if (i != j) D[i * n + j] = P[i] / (P[j] * (x[i] - x[j]));
  D[0] = -N * (N + 1.0) / 4.0;
  D[(size_t)N * n + N] = N * (N + 1.0) / 4.0;
}

bool basket_colouring(int64_t n_el, int n_loc, const int32_t* nodes, int64_t n_nodes,
                      std::vector<int8_t>& basket, std::vector<int32_t>& row, int64_t& rows) {
  const int64_t L = n_el * n_loc;
  std::vector<int32_t> cnt(n_nodes + 1, 0);
  for (int64_t q = 0; q < L; ++q) cnt[nodes[q] + 1]++;
  for (int64_t k = 0; k < n_nodes; ++k) cnt[k + 1] += cnt[k];
  std::vector<int32_t> occ(L), fill(cnt.begin(), cnt.end() - 1);
  for (int64_t q = 0; q < L; ++q) occ[fill[nodes[q]]++] = (int32_t)q;  // element-major order
  basket.assign(L, -1);
  row.assign(L, -1);
  int64_t H[12] = {0};
  for (int64_t k = 0; k < n_nodes; ++k) {
    int c = cnt[k + 1] - cnt[k];
    if (c == 0) continue;
    if (c > 12) return false;
    int S[12];
    std::iota(S, S + 12, 0);
    std::stable_sort(S, S + 12, [&](int a, int b) { return H[a] < H[b]; });
    for (int i = 0; i < c; ++i) {
      int b = S[i];
      int32_t q = occ[cnt[k] + i];
      basket[q] = (int8_t)b;
      row[q] = (int32_t)H[b];
      H[b]++;
    }
  }
  rows = (L + 11) / 12;
  return true;
}

void inverse_map(int64_t n_el, int n_loc, const int32_t* nodes, int64_t n_nodes,
                 const std::vector<int8_t>& basket, std::vector<int32_t>& rowptr,
                 std::vector<int32_t>& occ) {
  const int64_t L = n_el * n_loc;
  rowptr.assign(n_nodes + 1, 0);
  for (int64_t q = 0; q < L; ++q) rowptr[nodes[q] + 1]++;
  for (int64_t k = 0; k < n_nodes; ++k) rowptr[k + 1] += rowptr[k];
  occ.assign(L, 0);
  std::vector<int32_t> fill(rowptr.begin(), rowptr.end() - 1);
  for (int64_t q = 0; q < L; ++q) occ[fill[nodes[q]]++] = (int32_t)q;
  for (int64_t k = 0; k < n_nodes; ++k)
    std::sort(occ.begin() + rowptr[k], occ.begin() + rowptr[k + 1],
              [&](int32_t a, int32_t b) { return basket[a] < basket[b]; });
}

static uint64_t spread_bits(uint64_t v) {  // 32 -> 64 bit interleave
  v &= 0xffffffffull;
  v = (v | (v << 16)) & 0x0000ffff0000ffffull;
  v = (v | (v << 8)) & 0x00ff00ff00ff00ffull;
  v = (v | (v << 4)) & 0x0f0f0f0f0f0f0f0full;
  v = (v | (v << 2)) & 0x3333333333333333ull;
  v = (v | (v << 1)) & 0x5555555555555555ull;
  return v;
}

static double median(std::vector<double> v) {
  if (v.empty()) return 0.0;
  std::nth_element(v.begin(), v.begin() + v.size() / 2, v.end());
  return v[v.size() / 2];
}

void even_odd_table(int n, const std::vector<double>& M, std::vector<double>& out) {
  const int h = n / 2;
  for (int i = 0; i < h; ++i)
    for (int m = 0; m < h; ++m) out.push_back(0.5 * (M[i * n + m] - M[i * n + n - 1 - m]));
  for (int i = 0; i < h; ++i)
    for (int m = 0; m < h; ++m) out.push_back(0.5 * (M[i * n + m] + M[i * n + n - 1 - m]));
  if (n % 2) {
    for (int i = 0; i < h; ++i) out.push_back(M[i * n + h]);
    for (int m = 0; m < h; ++m) out.push_back(M[h * n + m]);
  }
}

bool build_patches(int64_t n_el, int n_loc, const int32_t* nodes, int64_t n_nodes,
                   const double* coords, int P, const std::vector<int8_t>& basket,
                   PatchPlan& pl, const int32_t* group) {
  if (P <= 0 || (int64_t)P * n_loc > (int64_t)kOccMore - 1) return false;  // occurrence values below the marks
  // --- element order: Morton code of the (x, y) cell, layer as minor key
  std::vector<double> cx(n_el), cy(n_el), cz(n_el), ex(n_el), ey(n_el), ez(n_el);
  double xmin = 1e300, ymin = 1e300, zmin = 1e300;
  for (int64_t e = 0; e < n_el; ++e) {
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300}, s[3] = {0, 0, 0};
    for (int p = 0; p < n_loc; ++p) {
      const double* c = coords + 3 * (int64_t)nodes[e * n_loc + p];
      for (int d = 0; d < 3; ++d) { s[d] += c[d]; lo[d] = std::min(lo[d], c[d]); hi[d] = std::max(hi[d], c[d]); }
    }
    cx[e] = s[0] / n_loc; cy[e] = s[1] / n_loc; cz[e] = s[2] / n_loc;
    ex[e] = hi[0] - lo[0]; ey[e] = hi[1] - lo[1]; ez[e] = hi[2] - lo[2];
    xmin = std::min(xmin, lo[0]); ymin = std::min(ymin, lo[1]); zmin = std::min(zmin, lo[2]);
  }
  const double hx = std::max(median(ex), 1e-300), hy = std::max(median(ey), 1e-300), hz = median(ez);
  std::vector<uint64_t> kxy(n_el), kz(n_el);
  for (int64_t e = 0; e < n_el; ++e) {
    const uint64_t ix = (uint64_t)std::max(0.0, std::floor((cx[e] - xmin) / hx));
    const uint64_t iy = (uint64_t)std::max(0.0, std::floor((cy[e] - ymin) / hy));
    kxy[e] = spread_bits(ix) | (spread_bits(iy) << 1);
    kz[e] = hz > 0 ? (uint64_t)std::max(0.0, std::floor((cz[e] - zmin) / hz)) : 0;
  }
  pl.perm.resize(n_el);
  std::iota(pl.perm.begin(), pl.perm.end(), 0);
  std::stable_sort(pl.perm.begin(), pl.perm.end(), [&](int32_t a, int32_t b) {
    if (group && group[a] != group[b]) return group[a] < group[b];
    if (kxy[a] != kxy[b]) return kxy[a] < kxy[b];
    return kz[a] < kz[b];
  });
  // --- global occurrence counts
  std::vector<int32_t> cnt(n_nodes, 0);
  for (int64_t q = 0; q < n_el * n_loc; ++q) cnt[nodes[q]]++;
  pl.P = P;
  // chunks of P elements in the new order, cut at case boundaries
  pl.el0.assign(1, 0);
  for (int64_t e = 0; e < n_el;) {
    int64_t end = std::min<int64_t>(e + P, n_el);
    if (group)
      for (int64_t q = e + 1; q < end; ++q)
        if (group[pl.perm[q]] != group[pl.perm[e]]) { end = q; break; }
    pl.el0.push_back((int32_t)end);
    e = end;
  }
  pl.n_patch = (int64_t)pl.el0.size() - 1;
  pl.pn_base.assign(1, 0);
  pl.pn_owned.clear(); pl.pn_id.clear(); pl.slot_base.clear(); pl.occ_ptr.clear();
  pl.occ.clear(); pl.lmap.clear();
  pl.optr_base.clear(); pl.occ_base.clear(); pl.lm_base.clear();
  pl.occ8.clear(); pl.o8_base.clear();
  pl.n_templates = 0;
  pl.max_pn = 0;
  // templates: content hash -> ids; per id the offsets and lengths of its three blocks
  struct Tmpl { int32_t lm, lm_n, op, op_n, oc, oc_n, o8; };
  std::vector<Tmpl> tmpl;
  std::unordered_map<uint64_t, std::vector<int32_t>> tmpl_index;
  std::vector<uint16_t> t_lm, t_op, t_oc, t_o8;
  std::vector<std::pair<int32_t, int32_t>> shared;  // (node, slot) in patch order
  int64_t slot = 0;
  struct Occ { int32_t node; int8_t bk; uint16_t lq; };
  std::vector<Occ> oc;
  std::vector<int32_t> local(n_nodes, -1), inpatch(n_nodes, 0);
  for (int64_t b = 0; b < pl.n_patch; ++b) {
    oc.clear();
    for (int32_t e = pl.el0[b]; e < pl.el0[b + 1]; ++e) {
      const int32_t orig = pl.perm[e];
      const int le = e - pl.el0[b];
      for (int p = 0; p < n_loc; ++p) {
        const int64_t q = (int64_t)orig * n_loc + p;
        oc.push_back({nodes[q], basket[q], (uint16_t)(le * n_loc + p)});
      }
    }
    std::sort(oc.begin(), oc.end(), [](const Occ& x, const Occ& y) {
      if (x.node != y.node) return x.node < y.node;
      return x.bk < y.bk;
    });
    std::vector<int32_t> own, shr;
    for (size_t r = 0; r < oc.size(); ++r) {
      if (r == 0 || oc[r].node != oc[r - 1].node) {
        int32_t c = 0;
        for (size_t t = r; t < oc.size() && oc[t].node == oc[r].node; ++t) ++c;
        (c == cnt[oc[r].node] ? own : shr).push_back(oc[r].node);
      }
    }
    const int32_t pn = (int32_t)(own.size() + shr.size());
    pl.max_pn = std::max(pl.max_pn, (int)pn);
    pl.pn_owned.push_back((int32_t)own.size());
    pl.slot_base.push_back((int32_t)slot);
    int32_t ln = 0;
    for (auto* lst : {&own, &shr})
      for (int32_t nd : *lst) { local[nd] = ln++; pl.pn_id.push_back(nd); }
    for (size_t r = 0; r < shr.size(); ++r) shared.push_back({shr[r], (int32_t)(slot + r)});
    slot += (int64_t)shr.size();
    // occurrences grouped per local node in basket order
    std::vector<std::vector<uint16_t>> per(pn);
    for (const Occ& o : oc) per[local[o.node]].push_back(o.lq);
    t_op.assign(1, 0);
    t_oc.clear();
    uint16_t acc = 0;
    for (int32_t k = 0; k < pn; ++k) {
      for (uint16_t v : per[k]) t_oc.push_back(v);
      acc = (uint16_t)(acc + per[k].size());
      t_op.push_back(acc);
    }
    t_lm.assign((size_t)(pl.el0[b + 1] - pl.el0[b]) * n_loc, 0);
    for (const Occ& o : oc) t_lm[o.lq] = (uint16_t)local[o.node];
    // the patch's template: an existing one with the same three blocks, else a new one
    uint64_t h = 1469598103934665603ull;
    for (const auto* v : {&t_lm, &t_op, &t_oc}) {
      h = (h ^ v->size()) * 1099511628211ull;
      for (uint16_t x : *v) h = (h ^ x) * 1099511628211ull;
    }
    auto same = [](const std::vector<uint16_t>& v, const std::vector<uint16_t>& all, int32_t off, int32_t n) {
      return (int32_t)v.size() == n && std::equal(v.begin(), v.end(), all.begin() + off);
    };
    std::vector<int32_t>& cand = tmpl_index[h];
    int32_t tid = -1;
    for (int32_t t : cand)
      if (same(t_lm, pl.lmap, tmpl[t].lm, tmpl[t].lm_n) && same(t_op, pl.occ_ptr, tmpl[t].op, tmpl[t].op_n) &&
          same(t_oc, pl.occ, tmpl[t].oc, tmpl[t].oc_n)) {
        tid = t;
        break;
      }
    if (tid < 0) {
      tid = (int32_t)tmpl.size();
      tmpl.push_back({(int32_t)pl.lmap.size(), (int32_t)t_lm.size(), (int32_t)pl.occ_ptr.size(), (int32_t)t_op.size(),
                      (int32_t)pl.occ.size(), (int32_t)t_oc.size(), (int32_t)pl.occ8.size()});
      // the template's occurrence records: kOccRec entries per node, in list order, unused
      // entries kOccEnd; a node with more occurrences keeps the first kOccRec - 1 and kOccMore
      t_o8.assign((size_t)pn * kOccRec, kOccEnd);
      for (int32_t k = 0; k < pn; ++k) {
        const size_t c = per[k].size();
        const size_t nk = c <= (size_t)kOccRec ? c : (size_t)kOccRec - 1;
        for (size_t r = 0; r < nk; ++r) t_o8[(size_t)k * kOccRec + r] = per[k][r];
        if (c > (size_t)kOccRec) t_o8[(size_t)k * kOccRec + kOccRec - 1] = kOccMore;
      }
      pl.occ8.insert(pl.occ8.end(), t_o8.begin(), t_o8.end());
      pl.lmap.insert(pl.lmap.end(), t_lm.begin(), t_lm.end());
      pl.occ_ptr.insert(pl.occ_ptr.end(), t_op.begin(), t_op.end());
      pl.occ.insert(pl.occ.end(), t_oc.begin(), t_oc.end());
      cand.push_back(tid);
    }
    pl.lm_base.push_back(tmpl[tid].lm);
    pl.optr_base.push_back(tmpl[tid].op);
    pl.occ_base.push_back(tmpl[tid].oc);
    pl.o8_base.push_back(tmpl[tid].o8);
    pl.pn_base.push_back((int32_t)pl.pn_id.size());
    for (const Occ& o : oc) local[o.node] = -1;
  }
  pl.n_slots = slot;
  pl.n_templates = (int64_t)tmpl.size();
  // internal node numbering: patch by patch, local order, first occurrence wins (a patch's
  // owned nodes occur in no other patch, so they get consecutive ids); nodes that belong
  // to no element go last
  pl.node_new.assign(n_nodes, -1);
  int32_t next = 0;
  for (int32_t nd : pl.pn_id)
    if (pl.node_new[nd] < 0) pl.node_new[nd] = next++;
  for (int64_t k = 0; k < n_nodes; ++k)
    if (pl.node_new[k] < 0) pl.node_new[k] = next++;
  pl.node_old.assign(n_nodes, 0);
  for (int64_t k = 0; k < n_nodes; ++k) pl.node_old[pl.node_new[k]] = (int32_t)k;
  for (auto& nd : pl.pn_id) nd = pl.node_new[nd];
  for (auto& pr : shared) pr.first = pl.node_new[pr.first];
  // boundary nodes in internal order, each with its slots in patch order
  std::stable_sort(shared.begin(), shared.end(),
                   [](const std::pair<int32_t, int32_t>& x, const std::pair<int32_t, int32_t>& y) {
                     return x.first < y.first;
                   });
  pl.sh_id.clear(); pl.sh_ptr.assign(1, 0); pl.sh_slot.clear();
  for (size_t r = 0; r < shared.size(); ++r) {
    if (r == 0 || shared[r].first != shared[r - 1].first) {
      if (r) pl.sh_ptr.push_back((int32_t)pl.sh_slot.size());
      pl.sh_id.push_back(shared[r].first);
    }
    pl.sh_slot.push_back(shared[r].second);
  }
  if (!shared.empty()) pl.sh_ptr.push_back((int32_t)pl.sh_slot.size());
  return true;
}

// Lagrange polynomials on the nodes x and their derivatives at t.
static void lagrange_at(const std::vector<double>& x, double t, double* L, double* dL) {
  const int n = (int)x.size();
  for (int m = 0; m < n; ++m) {
    double v = 1.0;
    for (int q = 0; q < n; ++q)
      if (q != m) v *= (t - x[q]) / (x[m] - x[q]);
    L[m] = v;
    if (dL) {
      double d = 0.0;
      for (int sk = 0; sk < n; ++sk) {
        if (sk == m) continue;
        double term = 1.0 / (x[m] - x[sk]);
        for (int q = 0; q < n; ++q)
          if (q != m && q != sk) term *= (t - x[q]) / (x[m] - x[q]);
        d += term;
      }
      dL[m] = d;
    }
  }
}

bool grid_plan(int64_t n_el, int n, const int32_t* nodes, int64_t n_nodes, const double* coords,
               int64_t n_search, int Nx, int Ny, GridPlan& g, std::string& err) {
  constexpr double kEps = 1e-8;  // spec2meshgrid_flat_shell.m:58
  const int nl = n * n;
  if (n_search <= 0 || n_search > n_el) { err = "n_search out of range"; return false; }
  double xmn = coords[0], xmx = coords[0], ymn = coords[1], ymx = coords[1];
  for (int64_t k = 1; k < n_nodes; ++k) {
    xmn = std::min(xmn, coords[3 * k]); xmx = std::max(xmx, coords[3 * k]);
    ymn = std::min(ymn, coords[3 * k + 1]); ymx = std::max(ymx, coords[3 * k + 1]);
  }
  g.Nx = Nx; g.Ny = Ny; g.n = n;
  const double L = xmx - xmn - 2 * kEps, B = ymx - ymn - 2 * kEps;
  g.xi.resize(Nx); g.yi.resize(Ny);
  for (int c = 0; c < Nx; ++c) g.xi[c] = xmn + kEps + L / (Nx - 1) * c;   // :103-106
  for (int r = 0; r < Ny; ++r) g.yi[r] = ymn + kEps + B / (Ny - 1) * r;
  const int64_t P = (int64_t)Nx * Ny;
  // element corners, counter-clockwise (:73), edge terms and bounding boxes (:121-170)
  const int mc[4] = {0, n - 1, nl - 1, nl - n};
  std::vector<double> cx(4 * n_search), cy(4 * n_search), bb(4 * n_search);
  for (int64_t e = 0; e < n_search; ++e) {
    for (int q = 0; q < 4; ++q) {
      const int64_t nd = nodes[e * nl + mc[q]];
      cx[4 * e + q] = coords[3 * nd];
      cy[4 * e + q] = coords[3 * nd + 1];
    }
    bb[4 * e] = *std::min_element(&cx[4 * e], &cx[4 * e] + 4);
    bb[4 * e + 1] = *std::min_element(&cy[4 * e], &cy[4 * e] + 4);
    bb[4 * e + 2] = *std::max_element(&cx[4 * e], &cx[4 * e] + 4);
    bb[4 * e + 3] = *std::max_element(&cy[4 * e], &cy[4 * e] + 4);
  }
  // buckets: every element in each bucket its box overlaps, in element order, so that a
  // point's candidates (box holds the point) are exactly those of a full scan, in order
  const int nb = std::max(1, (int)std::sqrt((double)n_search));
  const double bw = (xmx - xmn) / nb, bh = (ymx - ymn) / nb;
  auto bucket = [&](double v, double lo, double w) {
    int k = w > 0 ? (int)std::floor((v - lo) / w) : 0;
    return std::min(std::max(k, 0), nb - 1);
  };
  std::vector<std::vector<int32_t>> cell((size_t)nb * nb);
  for (int64_t e = 0; e < n_search; ++e)
    for (int by = bucket(bb[4 * e + 1], ymn, bh); by <= bucket(bb[4 * e + 3], ymn, bh); ++by)
      for (int bx = bucket(bb[4 * e], xmn, bw); bx <= bucket(bb[4 * e + 2], xmn, bw); ++bx)
        cell[(size_t)by * nb + bx].push_back((int32_t)e);
  g.owner.assign(P, -1);
  for (int c = 0; c < Nx; ++c)
    for (int r = 0; r < Ny; ++r) {
      const int64_t j = r + (int64_t)Ny * c;
      const double xp = g.xi[c], yp = g.yi[r];
      const auto& cand = cell[(size_t)bucket(yp, ymn, bh) * nb + bucket(xp, xmn, bw)];
      int32_t first = -1, inside = -1, count = 0;
      for (int32_t e : cand) {
        if (!(xp >= bb[4 * e] && xp <= bb[4 * e + 2] && yp >= bb[4 * e + 1] && yp <= bb[4 * e + 3])) continue;
        if (first < 0) first = e;
        ++count;
        if (inside < 0) {
          const double* X = &cx[4 * e];
          const double* Y = &cy[4 * e];
          bool neg = true;
          for (int q = 0; q < 4 && neg; ++q) {
            const int q1 = (q + 1) % 4;
            const double dx = X[q1] - X[q], dy = Y[q1] - Y[q];
            // (xp*y21 - x1*y21) - (yp*x21 - y1*x21) as written at :175-178
            const double d = (xp * dy - X[q] * dy) - (yp * dx - Y[q] * dx);
            neg = d < 0;
          }
          if (neg) inside = e;
        }
      }
      if (first < 0) {
        err = "grid point (" + std::to_string(xp) + ", " + std::to_string(yp) + ") lies in no element";
        return false;
      }
      g.owner[j] = count == 1 ? first : (inside >= 0 ? inside : first);
    }
  // local coordinates (:188-299)
  std::vector<double> x, w;
  gll(n, x, w);
  g.s.resize(P); g.t.resize(P); g.Lx.resize(P * n); g.Ly.resize(P * n);
  std::vector<double> Ls(n), Lt(n), dLs(n), dLt(n);
  for (int64_t j = 0; j < P; ++j) {
    const double xp = g.xi[j / Ny], yp = g.yi[j % Ny];
    const int32_t* en = nodes + (int64_t)g.owner[j] * nl;
    int k0 = 0;
    double best = 0.0;
    for (int k = 0; k < nl; ++k) {
      const double d = std::hypot(coords[3 * en[k]] - xp, coords[3 * en[k] + 1] - yp);
      if (k == 0 || d < best) { best = d; k0 = k; }
    }
    double s = x[k0 % n], t = x[k0 / n];
    double px = coords[3 * en[k0]], py = coords[3 * en[k0] + 1];
    auto inv_jac = [&](double ss, double tt, double* iJ) {
      lagrange_at(x, ss, Ls.data(), dLs.data());
      lagrange_at(x, tt, Lt.data(), dLt.data());
      double J11 = 0, J12 = 0, J21 = 0, J22 = 0;
      for (int b = 0; b < n; ++b)
        for (int a = 0; a < n; ++a) {
          const int64_t nd = en[b * n + a];
          const double Nx_ = dLs[a] * Lt[b], Ny_ = Ls[a] * dLt[b];
          J11 += Nx_ * coords[3 * nd]; J12 += Ny_ * coords[3 * nd];
          J21 += Nx_ * coords[3 * nd + 1]; J22 += Ny_ * coords[3 * nd + 1];
        }
      const double aa = 1.0 / (J11 * J22 - J12 * J21);  // inv_jacp_2D.m
      iJ[0] = aa * J22; iJ[1] = aa * (-J12); iJ[2] = aa * (-J21); iJ[3] = aa * J11;
    };
    auto position = [&](double ss, double tt) {
      lagrange_at(x, ss, Ls.data(), nullptr);
      lagrange_at(x, tt, Lt.data(), nullptr);
      px = py = 0.0;
      for (int b = 0; b < n; ++b)
        for (int a = 0; a < n; ++a) {
          const int64_t nd = en[b * n + a];
          px += Ls[a] * Lt[b] * coords[3 * nd];
          py += Ls[a] * Lt[b] * coords[3 * nd + 1];
        }
    };
    double iJ[4];
    inv_jac(x[0], x[0], iJ);  // the element's first node (:214-217)
    double s1 = s + iJ[0] * (xp - px) + iJ[1] * (yp - py);
    double t1 = t + iJ[2] * (xp - px) + iJ[3] * (yp - py);
    s = s1; t = t1;
    position(s, t);
    for (int it = 0; it < 3; ++it) {
      inv_jac(s, t, iJ);
      s1 = s + iJ[0] * (xp - px) + iJ[1] * (yp - py);
      t1 = t + iJ[2] * (xp - px) + iJ[3] * (yp - py);
      s = s1; t = t1;
      position(s, t);
    }
    g.s[j] = s; g.t[j] = t;
    lagrange_at(x, s, &g.Lx[j * n], nullptr);
    lagrange_at(x, t, &g.Ly[j * n], nullptr);
  }
  return true;
}

}  // namespace semhost

using namespace semhost;

// ---------------------------------------------------------------------------
// Mesh front end
// ---------------------------------------------------------------------------
// Local indices of the four edges of an n x n element (xi fastest), each from one corner
// to the next, corners included.
static void element_edges(int n, std::vector<std::vector<int>>& ed) {
  ed.assign(4, {});
  for (int k = 0; k < n; ++k) {
    ed[0].push_back(k);                  // eta = -1
    ed[1].push_back(k * n + n - 1);      // xi = +1
    ed[2].push_back((n - 1) * n + k);    // eta = +1
    ed[3].push_back(k * n);              // xi = -1
  }
}

extern "C" int sem_grid_plan(int64_t n_el, int n_xy, const int32_t* nodes, int64_t n_nodes, const double* coords,
                             int64_t n_search, int Nx, int Ny, double* xi, double* yi, int32_t* owner, double* s,
                             double* t) {
  // the host half of spec2meshgrid_flat_shell.m (:103-299), without a device
  if (n_el <= 0 || n_xy < 2 || n_xy > 8 || !nodes || n_nodes <= 0 || !coords || Nx < 2 || Ny < 2) return SEM_E_ARG;
  for (int64_t q = 0; q < n_el * n_xy * n_xy; ++q)
    if (nodes[q] < 0 || nodes[q] >= n_nodes) return SEM_E_ARG;
  GridPlan g;
  std::string err;
  if (!grid_plan(n_el, n_xy, nodes, n_nodes, coords, n_search > 0 ? n_search : n_el, Nx, Ny, g, err))
    return SEM_E_ARG;
  if (xi) std::copy(g.xi.begin(), g.xi.end(), xi);
  if (yi) std::copy(g.yi.begin(), g.yi.end(), yi);
  if (owner) std::copy(g.owner.begin(), g.owner.end(), owner);
  if (s) std::copy(g.s.begin(), g.s.end(), s);
  if (t) std::copy(g.t.begin(), g.t.end(), t);
  return SEM_OK;
}

extern "C" int sem_mesh_boundary_nodes(int64_t n_el, int n_xy, const int32_t* nodes, int32_t* bnd,
                                       int64_t* n_bnd) {
  // quad2spectral_Fiborek.m:153-156: nodes of the element edges that belong to one
  // element (edges keyed by their corner nodes, as the reference's unique edge midpoints)
  if (n_el <= 0 || n_xy < 2 || !nodes || !bnd || !n_bnd) return SEM_E_ARG;
  const int nl = n_xy * n_xy;
  std::vector<std::vector<int>> ed;
  element_edges(n_xy, ed);
  auto key = [&](int64_t e, int s) {
    const int64_t a = nodes[e * nl + ed[s].front()], b = nodes[e * nl + ed[s].back()];
    return std::make_pair(std::min(a, b), std::max(a, b));
  };
  std::vector<std::pair<int64_t, int64_t>> keys;
  keys.reserve(n_el * 4);
  for (int64_t e = 0; e < n_el; ++e)
    for (int s = 0; s < 4; ++s) keys.push_back(key(e, s));
  std::vector<std::pair<int64_t, int64_t>> sorted = keys;
  std::sort(sorted.begin(), sorted.end());
  auto count = [&](const std::pair<int64_t, int64_t>& k) {
    const auto r = std::equal_range(sorted.begin(), sorted.end(), k);
    return r.second - r.first;
  };
  std::vector<int32_t> out;
  for (int64_t e = 0; e < n_el; ++e)
    for (int s = 0; s < 4; ++s)
      if (count(keys[e * 4 + s]) == 1)
        for (int l : ed[s]) out.push_back(nodes[e * nl + l]);
  std::sort(out.begin(), out.end());
  out.erase(std::unique(out.begin(), out.end()), out.end());
  std::copy(out.begin(), out.end(), bnd);
  *n_bnd = (int64_t)out.size();
  return SEM_OK;
}

extern "C" int sem_mesh_split_delam(int64_t n_el, int n_xy, const int32_t* nodes, int64_t n_nodes,
                                    const double* coords, int64_t n_den, const int32_t* den, int64_t n_bnd,
                                    const int32_t* bnd, int32_t* nodes_out, double* coords_out,
                                    int64_t* n_nodes_out, int32_t* den_under, int32_t* den_above) {
  // split_delam_nodes_flat_shell.m:1-276.  The delaminated elements are appended again
  // (the copies, den_above, lie above the split interface).  A corner of a copy is a rim
  // node when at most one other delaminated element has it as a corner (:68-88), an
  // edge node when no other delaminated element has it on an edge (:90-236).  Rim nodes
  // on the plate boundary are split too, except the "free" corners (no delaminated
  // neighbour) there, of which the smallest is dropped when there are three (:238-246).
  // Every other node of the copies gets a new id, in ascending order of the old ids
  // (:247-260), with the coordinates of its twin (:263-266).
  if (n_el <= 0 || n_xy < 2 || !nodes || !coords || n_den < 0 || (n_den && !den) || n_bnd < 0 ||
      (n_bnd && !bnd) || !nodes_out || !coords_out || !n_nodes_out || (n_den && (!den_under || !den_above)))
    return SEM_E_ARG;
  const int nl = n_xy * n_xy;
  for (int64_t k = 0; k < n_den; ++k)
    if (den[k] < 0 || den[k] >= n_el) return SEM_E_ARG;
  int64_t nodemax = -1;
  for (int64_t q = 0; q < n_el * nl; ++q) {
    if (nodes[q] < 0 || nodes[q] >= n_nodes) return SEM_E_ARG;
    nodemax = std::max<int64_t>(nodemax, nodes[q]);
  }
  const int64_t nm = nodemax;
  std::copy(nodes, nodes + n_el * nl, nodes_out);
  for (int64_t k = 0; k < n_den; ++k) {
    std::copy(nodes + (int64_t)den[k] * nl, nodes + ((int64_t)den[k] + 1) * nl, nodes_out + (n_el + k) * nl);
    den_under[k] = den[k];
    den_above[k] = (int32_t)(n_el + k);
  }
  const int mc[4] = {0, n_xy - 1, nl - n_xy, nl - 1};
  std::vector<int> edge_loc;  // edge-interior local indices of all four edges
  {
    std::vector<std::vector<int>> ed;
    element_edges(n_xy, ed);
    for (const auto& e : ed)
      for (size_t t = 1; t + 1 < e.size(); ++t) edge_loc.push_back(e[t]);
  }
  // occurrences as a corner / on an edge over all delaminated elements; an element's
  // "neighbours" count is that total less its own occurrence
  std::unordered_map<int64_t, int> corner_cnt, edge_cnt;
  for (int64_t k = 0; k < n_den; ++k) {
    const int32_t* row = nodes + (int64_t)den[k] * nl;
    for (int c : mc) ++corner_cnt[row[c]];
    for (int l : edge_loc) ++edge_cnt[row[l]];
  }
  std::vector<int64_t> outer, free_corner;
  for (int64_t k = 0; k < n_den; ++k) {
    const int32_t* row = nodes + (int64_t)den[k] * nl;
    for (int c : mc) {
      const int nb = corner_cnt[row[c]] - 1;
      if (nb <= 1) outer.push_back(row[c]);
      if (nb == 0) free_corner.push_back(row[c]);
    }
    for (int l : edge_loc)
      if (edge_cnt[row[l]] - 1 == 0) outer.push_back(row[l]);
  }
  std::vector<int64_t> bset(bnd, bnd + n_bnd);
  std::sort(bset.begin(), bset.end());
  bset.erase(std::unique(bset.begin(), bset.end()), bset.end());
  auto uniq = [](std::vector<int64_t>& v) {
    std::sort(v.begin(), v.end());
    v.erase(std::unique(v.begin(), v.end()), v.end());
  };
  uniq(free_corner);
  std::vector<int64_t> special;
  std::set_intersection(free_corner.begin(), free_corner.end(), bset.begin(), bset.end(),
                        std::back_inserter(special));
  if (special.size() == 3) special.erase(special.begin());
  uniq(outer);
  std::vector<int64_t> rim;
  std::set_difference(outer.begin(), outer.end(), bset.begin(), bset.end(), std::back_inserter(rim));
  rim.insert(rim.end(), special.begin(), special.end());
  uniq(rim);
  std::vector<int64_t> delam;
  for (int64_t k = 0; k < n_den; ++k)
    for (int l = 0; l < nl; ++l) delam.push_back(nodes[(int64_t)den[k] * nl + l]);
  uniq(delam);
  std::vector<int64_t> inner;
  std::set_difference(delam.begin(), delam.end(), rim.begin(), rim.end(), std::back_inserter(inner));
  if (nodemax + (int64_t)inner.size() >= (int64_t)INT32_MAX) return SEM_E_ARG;
  std::unordered_map<int64_t, int32_t> renum;
  for (int64_t nd : inner) renum[nd] = (int32_t)(++nodemax);
  for (int64_t k = 0; k < n_den; ++k) {
    int32_t* row = nodes_out + (n_el + k) * nl;
    for (int l = 0; l < nl; ++l) {
      const auto it = renum.find(row[l]);
      if (it != renum.end()) row[l] = it->second;
    }
  }
  const int64_t n_new = nodemax + 1;
  std::fill(coords_out, coords_out + 3 * n_new, 0.0);
  std::copy(coords, coords + 3 * (nm + 1), coords_out);
  for (int64_t k = 0; k < n_den; ++k)
    for (int l = 0; l < nl; ++l) {
      const int64_t a = nodes_out[(n_el + k) * nl + l], u = nodes_out[(int64_t)den[k] * nl + l];
      for (int d = 0; d < 3; ++d) coords_out[3 * a + d] = coords_out[3 * u + d];
    }
  *n_nodes_out = n_new;
  return SEM_OK;
}

extern "C" int sem_gll(int n, double* x, double* w) {
  // gll.m:1-39 (the reference's table for n <= 10, the closed form beyond; see semhost::gll)
  if (n < 2 || n > 64 || !x || !w) return SEM_E_ARG;
  std::vector<double> xv, wv;
  gll(n, xv, wv);
  std::copy(xv.begin(), xv.end(), x);
  std::copy(wv.begin(), wv.end(), w);
  return SEM_OK;
}

static double matlab_round(double v) { return std::round(v); }  // half away from zero

extern "C" int sem_mesh_change_turn_quad(int64_t n_quads, int32_t* quads, int64_t n_pos, const double* pos,
                                         int64_t* n_flipped) {
  // change_turn_quad.m:1-110: a bilinear quad whose Jacobian determinant is not positive at
  // one of its corners is numbered the other way round (corner order 1, 4, 3, 2).  At a
  // corner the determinant is (a quarter of) the cross product of its two edges, taken
  // towards the next and the previous corner.
  if (n_quads < 0 || !quads || !pos || n_pos <= 0) return SEM_E_ARG;
  for (int64_t q = 0; q < 4 * n_quads; ++q)
    if (quads[q] < 0 || quads[q] >= n_pos) return SEM_E_ARG;
  static const int corner[4][3] = {{0, 1, 3}, {1, 2, 0}, {2, 3, 1}, {3, 0, 2}};
  int64_t flips = 0;
  for (int64_t e = 0; e < n_quads; ++e) {
    int32_t* qd = quads + 4 * e;
    bool bad = false;
    for (const auto& c : corner) {
      const double* a = pos + 3 * (int64_t)qd[c[0]];
      const double* b = pos + 3 * (int64_t)qd[c[1]];
      const double* d = pos + 3 * (int64_t)qd[c[2]];
      const double det = (b[0] - a[0]) * (d[1] - a[1]) - (b[1] - a[1]) * (d[0] - a[0]);
      if (det <= 0) { bad = true; break; }
    }
    if (bad) {
      std::swap(qd[1], qd[3]);
      ++flips;
    }
  }
  if (n_flipped) *n_flipped = flips;
  return SEM_OK;
}

extern "C" int sem_mesh_quad2spectral(int64_t n_quads, const int32_t* quads, int64_t n_fem,
                                      const double* pos, int order, int32_t* nodes_out,
                                      double* coords_out, int64_t* n_nodes) {
  // quad2spectral_Fiborek.m:1-173 (structure and arithmetic kept; own code).
  const int n = order + 1;
  if (n < 3 || n > 10 || n_quads <= 0 || !quads || !pos || !nodes_out || !coords_out || !n_nodes)
    return SEM_E_ARG;
  const int ni = n - 2;
  std::vector<double> ksi, wq;
  gll(n, ksi, wq);
  for (int64_t i = 0; i < 4 * n_quads; ++i)
    if (quads[i] < 0 || quads[i] >= n_fem) return SEM_E_ARG;
  // edge-interior node coordinates, order (element, edge, k) (:19-39)
  const int64_t ne = n_quads * 4 * ni;
  std::vector<std::array<double, 3>> E(ne);
  for (int64_t e = 0; e < n_quads; ++e) {
    for (int ed = 0; ed < 4; ++ed) {
      const double* a = pos + 3 * (int64_t)quads[4 * e + ed];
      const double* b = pos + 3 * (int64_t)quads[4 * e + (ed + 1) % 4];
      for (int d = 0; d < 3; ++d) {
        double mid = (b[d] + a[d]) / 2;
        mid = matlab_round(mid * 1e6) * 1e-6;
        double ln = (b[d] - a[d]) / 2;
        for (int k = 0; k < ni; ++k) {
          double v = mid + ln * ksi[k + 1];
          E[(e * 4 + ed) * ni + k][d] = v + 0.0;  // -0 -> +0
        }
      }
    }
  }
  // unique(...,'rows') (:41): lexicographic order of (x, y, z)
  std::vector<int64_t> perm(ne);
  std::iota(perm.begin(), perm.end(), 0);
  std::sort(perm.begin(), perm.end(), [&](int64_t a, int64_t b) {
    if (E[a][0] != E[b][0]) return E[a][0] < E[b][0];
    if (E[a][1] != E[b][1]) return E[a][1] < E[b][1];
    if (E[a][2] != E[b][2]) return E[a][2] < E[b][2];
    return a < b;
  });
  std::vector<int64_t> inv(ne);
  std::vector<std::array<double, 3>> U;
  U.reserve(ne / 2 + 1);
  for (int64_t r = 0; r < ne; ++r) {
    int64_t p = perm[r];
    if (r == 0 || E[p] != U.back()) U.push_back(E[p]);
    inv[p] = (int64_t)U.size() - 1;
  }
  const int64_t n_edge = (int64_t)U.size();
  const int64_t nint = (int64_t)ni * ni;
  const int64_t ntot = n_fem + n_edge + n_quads * nint;
  // coordinates: corners, unique edge nodes, interior nodes
  for (int64_t i = 0; i < 3 * n_fem; ++i) coords_out[i] = pos[i];
  for (int64_t i = 0; i < n_edge; ++i)
    for (int d = 0; d < 3; ++d) coords_out[3 * (n_fem + i) + d] = U[i][d];
  double zval = n_edge ? U[0][2] : 0.0;
  for (int64_t i = 0; i < n_edge; ++i)
    if (U[i][2] != zval) return SEM_E_UNSUPPORTED;  // flat meshes only (:102)
  std::vector<int64_t> e1(ni), e2(ni), e3(ni), e4(ni);
  for (int64_t e = 0; e < n_quads; ++e) {
    const int64_t* ed = inv.data() + e * 4 * ni;
    for (int k = 0; k < ni; ++k) {
      e1[k] = ed[k];
      e2[k] = ed[ni + k];
      e3[k] = ed[3 * ni - 1 - k];  // edge 3 reversed (:53)
      e4[k] = ed[4 * ni - 1 - k];  // edge 4 reversed (:54)
    }
    for (int q = 0; q < ni; ++q) {
      const double x2 = U[e2[q]][0], y2 = U[e2[q]][1], x4 = U[e4[q]][0], y4 = U[e4[q]][1];
      const double a42 = y2 - y4, b42 = x4 - x2;
      const double c42 = (y4 - y2) * x4 + (x2 - x4) * y4;
      for (int p = 0; p < ni; ++p) {
        const double x1 = U[e1[p]][0], y1 = U[e1[p]][1], x3 = U[e3[p]][0], y3 = U[e3[p]][1];
        const double a13 = y3 - y1, b13 = x1 - x3;
        const double c13 = (y1 - y3) * x1 + (x3 - x1) * y1;
        const double W = a13 * b42 - b13 * a42;
        const double Wx = b13 * c42 - c13 * b42;
        const double Wy = c13 * a42 - a13 * c42;
        int64_t id = n_fem + n_edge + e * nint + (int64_t)q * ni + p;
        coords_out[3 * id + 0] = Wx / W;
        coords_out[3 * id + 1] = Wy / W;
        coords_out[3 * id + 2] = zval;
      }
    }
    // local xi-fastest numbering (elementNodes_pl3, :104-118)
    int32_t* out = nodes_out + e * (int64_t)n * n;
    auto edge = [&](int which, int k) { return (int32_t)(n_fem + inv[e * 4 * ni + which * ni + k]); };
    auto inner = [&](int q, int p) { return (int32_t)(n_fem + n_edge + e * nint + (int64_t)q * ni + p); };
    int c = 0;
    out[c++] = quads[4 * e + 0];
    for (int k = 0; k < ni; ++k) out[c++] = edge(0, k);
    out[c++] = quads[4 * e + 1];
    for (int r = 0; r < ni; ++r) {
      out[c++] = edge(3, ni - 1 - r);
      for (int p = 0; p < ni; ++p) out[c++] = inner(r, p);
      out[c++] = edge(1, r);
    }
    out[c++] = quads[4 * e + 3];
    for (int k = ni - 1; k >= 0; --k) out[c++] = edge(2, k);
    out[c++] = quads[4 * e + 2];
  }
  *n_nodes = ntot;
  return SEM_OK;
}

extern "C" int sem_mesh_extrude(int64_t n_el2d, int n_xy, const int32_t* nodes2d,
                                int64_t n_nodes2d, const double* coords2d, int nz, int Nz,
                                const double* h, int32_t* nodes3d_out, double* coords3d_out) {
  // extrude_mesh.m:1-85 (0-based ids; keeps the hlay = dz update of :84).
  if (n_el2d <= 0 || n_xy < 2 || nz < 2 || Nz < 1 || !nodes2d || !coords2d || !h ||
      !nodes3d_out || !coords3d_out)
    return SEM_E_ARG;
  std::vector<double> ksi, wq;
  gll(nz, ksi, wq);
  const int nl2 = n_xy * n_xy, nl3 = nl2 * nz;
  const int64_t dofl = n_nodes2d;
  for (int i = 0; i < Nz; ++i)
    for (int64_t e = 0; e < n_el2d; ++e)
      for (int k = 0; k < nz; ++k) {
        int64_t lev = (int64_t)i * (nz - 1) + k;
        for (int c = 0; c < nl2; ++c)
          nodes3d_out[((int64_t)i * n_el2d + e) * nl3 + k * nl2 + c] =
              (int32_t)(nodes2d[e * nl2 + c] + dofl * lev);
      }
  double hlay = 0.0;
  for (int iz = 0; iz < Nz; ++iz) {
    double dz = h[iz];
    for (int k = 0; k < nz; ++k) {
      double zz = dz * (ksi[k] + 1) / 2 + hlay;
      int64_t lev = (int64_t)iz * (nz - 1) + k;
      for (int64_t v = 0; v < n_nodes2d; ++v) {
        int64_t id = v + dofl * lev;
        coords3d_out[3 * id + 0] = coords2d[3 * v + 0];
        coords3d_out[3 * id + 1] = coords2d[3 * v + 1];
        coords3d_out[3 * id + 2] = zz;
      }
    }
    hlay = dz;
  }
  return SEM_OK;
}

extern "C" int sem_mesh_baskets(int64_t n_el, int n_loc, const int32_t* nodes, int64_t n_nodes,
                                int32_t* I_G, int32_t* I_L) {
  if (n_el <= 0 || n_loc <= 0 || !nodes || !I_G || !I_L) return SEM_E_ARG;
  const int64_t L = n_el * n_loc;
  for (int64_t q = 0; q < L; ++q)
    if (nodes[q] < 0 || nodes[q] >= n_nodes) return SEM_E_ARG;
  std::vector<int8_t> basket;
  std::vector<int32_t> row;
  int64_t rows = 0;
  if (!basket_colouring(n_el, n_loc, nodes, n_nodes, basket, row, rows)) return SEM_E_ARG;
  std::fill(I_G, I_G + 12 * rows, 0);
  std::fill(I_L, I_L + 12 * rows, 0);
  for (int64_t q = 0; q < L; ++q) {
    const int64_t at = (int64_t)basket[q] * rows + row[q];
    I_G[at] = nodes[q] + 1;
    I_L[at] = (int32_t)(q + 1);
  }
  return SEM_OK;
}

// ---------------------------------------------------------------------------
// Laminate helpers
// ---------------------------------------------------------------------------
static void ply_bounds(int lay, const double* lh, double& h, std::vector<double>& h1,
                       std::vector<double>& h2) {
  // force_excit_c.m:75-83 -- ply 1 on top
  h = 0;
  for (int i = 0; i < lay; ++i) h += lh[i];
  h1.assign(lay, 0.0);
  h2.assign(lay, 0.0);
  h2[0] = h / 2;
  h1[0] = h / 2 - lh[0];
  for (int i = 1; i < lay; ++i) {
    h2[i] = h2[i - 1] - lh[i - 1];
    h1[i] = h1[i - 1] - lh[i];
  }
}

// composite_plate_c_delam.m:10-19 (rotation) and the through-thickness integrals over
// plies [i0, i1) at their z in the whole laminate, with the whole thickness h in the
// shear factor (:48-87 for all plies, :89-131 above a delamination, :133-173 under it)
static void laminate_shell_plies(int lay, const double* lh, const double* alpha_deg, const double* rho,
                                 const double* q, int i0, int i1, double* c) {
  double h;
  std::vector<double> h1, h2;
  ply_bounds(lay, lh, h, h1, h2);
  for (int k = 0; k < 23; ++k) c[k] = 0.0;
  for (int i = i0; i < i1; ++i) {
    const double* qi = q + 6 * i;
    const double q11 = qi[0], q12 = qi[1], q22 = qi[2], q44 = qi[3], q55 = qi[4], q66 = qi[5];
    const double al = alpha_deg[i] * M_PI / 180, m = std::cos(al), n = std::sin(al);
    const double m2 = m * m, n2 = n * n;
    const double s11 = q11 * m2 * m2 + 2 * (q12 + 2 * q66) * m2 * n2 + q22 * n2 * n2;
    const double s12 = (q11 + q22 - 4 * q66) * m2 * n2 + q12 * (m2 * m2 + n2 * n2);
    const double s16 = (q11 - q12 - 2 * q66) * m2 * m * n + (q12 - q22 + 2 * q66) * m * n2 * n;
    const double s22 = q11 * n2 * n2 + 2 * (q12 + 2 * q66) * m2 * n2 + q22 * m2 * m2;
    const double s26 = (q11 - q12 - 2 * q66) * n2 * n * m + (q12 - q22 + 2 * q66) * n * m2 * m;
    const double s44 = q44 * m2 + q55 * n2;
    const double s45 = (q55 - q44) * m * n;
    const double s55 = q55 * m2 + q44 * n2;
    const double s66 = (q11 + q22 - 2 * q12) * m2 * n2 + q66 * (m2 - n2) * (m2 - n2);
    const double t1 = h2[i] - h1[i];
    const double t2 = (h2[i] * h2[i] - h1[i] * h1[i]) / 2;
    const double t3 = (h2[i] * h2[i] * h2[i] - h1[i] * h1[i] * h1[i]) / 3;
    const double ts = 5.0 / 4.0 * (t1 - 4.0 / 3.0 * (h2[i] * h2[i] * h2[i] - h1[i] * h1[i] * h1[i]) / (h * h));
    const double s[6] = {s11, s12, s16, s22, s26, s66};
    for (int k = 0; k < 6; ++k) {
      c[k] += s[k] * t1;
      c[9 + k] += s[k] * t2;
      c[15 + k] += s[k] * t3;
    }
    c[6] += s44 * ts;
    c[7] += s45 * ts;
    c[8] += s55 * ts;
    c[21] += rho[i] * t1;
    c[22] += rho[i] * t3;
  }
}

extern "C" int sem_laminate_shell(int lay, const double* lh, const double* alpha_deg,
                                  const double* rho, const double* q, double* c) {
  if (lay <= 0 || !lh || !alpha_deg || !rho || !q || !c) return SEM_E_ARG;
  laminate_shell_plies(lay, lh, alpha_deg, rho, q, 0, lay, c);
  return SEM_OK;
}

extern "C" int sem_laminate_shell_delam(int lay, const double* lh, const double* alpha_deg,
                                        const double* rho, const double* q, int delamination_layer, int side,
                                        double* c) {
  // composite_plate_c_delam.m:89-131 (side > 0: plies 1..delamination_layer) and
  // :133-173 (side < 0: plies delamination_layer+1..lay)
  if (lay <= 0 || !lh || !alpha_deg || !rho || !q || !c || side == 0) return SEM_E_ARG;
  if (delamination_layer < 1 || delamination_layer >= lay) return SEM_E_ARG;
  if (side > 0) laminate_shell_plies(lay, lh, alpha_deg, rho, q, 0, delamination_layer, c);
  else laminate_shell_plies(lay, lh, alpha_deg, rho, q, delamination_layer, lay, c);
  return SEM_OK;
}

namespace {

// Rotated stiffness of one ply (composite3D_transf.m:66-77), as the 13 entries of the solid
// material row D11 D12 D13 D14 D22 D23 D24 D33 D34 D44 D55 D56 D66 with the reference's
// pairing (D55 = s55 with gamma_yz, D56 = s45, D66 = s44; :104-105, :116-117).
void ply_rotated_3d(const double* qi, double alpha_deg, double* d) {
  const double q11 = qi[0], q12 = qi[1], q13 = qi[2], q22 = qi[3], q23 = qi[4], q33 = qi[5],
               q44 = qi[6], q55 = qi[7], q66 = qi[8];
  const double eps = 2.220446049250313e-16;
  const double al = alpha_deg * M_PI / 180;
  double m = std::cos(al), n = std::sin(al);
  if (std::fabs(m) < eps) m = 0;  // :72-75
  if (std::fabs(n) < eps) n = 0;
  const double m2 = m * m, n2 = n * n;
  d[0] = q11 * m2 * m2 + 2 * (q12 + 2 * q66) * m2 * n2 + q22 * n2 * n2;                   // s11
  d[1] = (q11 + q22 - 4 * q66) * m2 * n2 + q12 * (m2 * m2 + n2 * n2);                      // s12
  d[2] = q13 * m2 + q23 * n2;                                                              // s13
  d[3] = (q11 - q12 - 2 * q66) * m2 * m * n + (q12 - q22 + 2 * q66) * m * n2 * n;          // s16
  d[4] = q11 * n2 * n2 + 2 * (q12 + 2 * q66) * m2 * n2 + q22 * m2 * m2;                   // s22
  d[5] = n2 * q13 + m2 * q23;                                                              // s23
  d[6] = (q11 - q12 - 2 * q66) * n2 * n * m + (q12 - q22 + 2 * q66) * n * m2 * m;          // s26
  d[7] = q33;                                                                              // s33
  d[8] = (q13 - q23) * m * n;                                                              // s36
  d[9] = (q11 + q22 - 2 * q12 - 2 * q66) * m2 * n2 + q66 * (m2 * m2 + n2 * n2);            // s66
  d[10] = q55 * m2 + q44 * n2;                                                             // s55
  d[11] = (q55 - q44) * m * n;                                                             // s45
  d[12] = q44 * m2 + q55 * n2;                                                             // s44
}

}  // namespace

extern "C" int sem_laminate_solid_smeared(int lay, const double* lh, const double* alpha_deg,
                                          const double* rho, const double* q, double* c) {
  // composite3D_transf.m:49-108, prop='smeared': thickness-weighted average of the
  // rotated ply stiffness, with D(5,5)=a55 (gamma_yz), D(5,6)=a45, D(6,6)=a44 (:104-105).
  if (lay <= 0 || !lh || !alpha_deg || !rho || !q || !c) return SEM_E_ARG;
  double h;
  std::vector<double> h1, h2;
  ply_bounds(lay, lh, h, h1, h2);
  double a[13] = {0}, J0 = 0;
  for (int i = 0; i < lay; ++i) {
    double d[13];
    ply_rotated_3d(q + 9 * i, alpha_deg[i], d);
    const double f = (h2[i] - h1[i]) / h;
    J0 += rho[i] * f;
    for (int k = 0; k < 13; ++k) a[k] += d[k] * f;
  }
  std::memcpy(c, a, sizeof(a));
  c[13] = J0;
  return SEM_OK;
}

extern "C" int sem_laminate_solid_layered(int lay, const double* alpha_deg, const double* rho, const double* q,
                                          double* c) {
  // composite3D_transf.m:109-119, prop='layered': one material row per ply, its rotated
  // stiffness (same entries and pairing as the smeared row) and its density (J0 = rho)
  if (lay <= 0 || !alpha_deg || !rho || !q || !c) return SEM_E_ARG;
  for (int i = 0; i < lay; ++i) {
    ply_rotated_3d(q + 9 * i, alpha_deg[i], c + 14 * i);
    c[14 * i + 13] = rho[i];
  }
  return SEM_OK;
}
