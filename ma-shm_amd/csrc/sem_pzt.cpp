// PZT transducers: piezoelectric constants, element coupling matrices, actuator
// equivalent forces and sensor voltage functionals.  Set-up work on the host; the time
// loop sees an actuator as point sources (sem_set_excitation) and a sensor as a linear
// output functional (sem_set_outputs), so the step kernels need nothing PZT-specific.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <vector>

#include "../../include/sem_hip.h"
#include "sem_host.hpp"

namespace {

using semhost::gll;
using semhost::lagrange_derivative;

// Dense LU with partial pivoting, A [n x n] row-major overwritten; solves A X = B for
// B [n x nrhs] row-major in place.  Returns false on a zero pivot.
bool lu_solve(int n, std::vector<double>& A, int nrhs, std::vector<double>& B) {
  std::vector<int> piv(n);
  for (int k = 0; k < n; ++k) {
    int p = k;
    double best = std::fabs(A[(size_t)k * n + k]);
    for (int r = k + 1; r < n; ++r)
      if (std::fabs(A[(size_t)r * n + k]) > best) { best = std::fabs(A[(size_t)r * n + k]); p = r; }
    if (best == 0.0) return false;
    if (p != k) {
      for (int c = 0; c < n; ++c) std::swap(A[(size_t)k * n + c], A[(size_t)p * n + c]);
      for (int c = 0; c < nrhs; ++c) std::swap(B[(size_t)k * nrhs + c], B[(size_t)p * nrhs + c]);
    }
    const double inv = 1.0 / A[(size_t)k * n + k];
    for (int r = k + 1; r < n; ++r) {
      const double f = A[(size_t)r * n + k] * inv;
      if (f == 0.0) continue;
      for (int c = k + 1; c < n; ++c) A[(size_t)r * n + c] -= f * A[(size_t)k * n + c];
      for (int c = 0; c < nrhs; ++c) B[(size_t)r * nrhs + c] -= f * B[(size_t)k * nrhs + c];
    }
  }
  for (int k = n - 1; k >= 0; --k) {
    const double inv = 1.0 / A[(size_t)k * n + k];
    for (int c = 0; c < nrhs; ++c) {
      double s = B[(size_t)k * nrhs + c];
      for (int j = k + 1; j < n; ++j) s -= A[(size_t)k * n + j] * B[(size_t)j * nrhs + c];
      B[(size_t)k * nrhs + c] = s * inv;
    }
  }
  return true;
}

bool inv6(const double* S, double* Q) {
  std::vector<double> A(S, S + 36), B(36, 0.0);
  for (int i = 0; i < 6; ++i) B[i * 7] = 1.0;
  if (!lu_solve(6, A, 6, B)) return false;
  std::copy(B.begin(), B.end(), Q);
  return true;
}

struct Gll {
  int n, nz;
  std::vector<double> x, w, D, xz, wz, Dz;
  Gll(int n_, int nz_) : n(n_), nz(nz_) {
    gll(n, x, w);
    lagrange_derivative(n, x, D);
    gll(nz, xz, wz);
    lagrange_derivative(nz, xz, Dz);
  }
};

// Element matrices of coupl3D_v2.m:1-160 (frame = 1: strains in the local frame v1 v2 v3
// built from the Jacobian rows) or coupl3D_global.m (frame = 0: strains in x y z; the
// electric gradient still uses v1 v2 v3, as coded there).  kuf [3*np x np], kff [np x np],
// row-major; e [3 x 6] row-major with columns in the strain order of B (ex ey ez gxy gyz gxz),
// g [3 x 3].  At a GLL point only the shape functions on the three lines through it have
// non-zero derivatives (the reference's 1e-12 zeroing of shape3D_prim_v2.m removes the
// round-off of the others).
void pzt_element(const Gll& t, const double* X, const double* e, const double* g, int frame,
                 std::vector<double>& kuf, std::vector<double>& kff) {
  const int n = t.n, nz = t.nz, np = n * n * nz;
  kuf.assign((size_t)3 * np * np, 0.0);
  kff.assign((size_t)np * np, 0.0);
  std::vector<int> S;
  std::vector<double> Bc, Bf;  // per support function: B [6 x 3], Bfi [3]
  for (int kz = 0; kz < nz; ++kz)
    for (int ky = 0; ky < n; ++ky)
      for (int kx = 0; kx < n; ++kx) {
        double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};  // J[r][c] = d x_c / d xi_r
        for (int m = 0; m < n; ++m) {
          const double* p1 = X + 3 * ((kz * n + ky) * n + m);
          const double* p2 = X + 3 * ((kz * n + m) * n + kx);
          for (int c = 0; c < 3; ++c) {
            J[0][c] += t.D[kx * n + m] * p1[c];
            J[1][c] += t.D[ky * n + m] * p2[c];
          }
        }
        for (int m = 0; m < nz; ++m) {
          const double* p3 = X + 3 * ((m * n + ky) * n + kx);
          for (int c = 0; c < 3; ++c) J[2][c] += t.Dz[kz * nz + m] * p3[c];
        }
        // local frame: v3 = normal (J row 1 x J row 2), v1 = J row 1, v2 = v3 x v1
        double v3[3] = {J[0][1] * J[1][2] - J[1][1] * J[0][2], J[1][0] * J[0][2] - J[0][0] * J[1][2],
                        J[0][0] * J[1][1] - J[1][0] * J[0][1]};
        double s = std::sqrt(v3[0] * v3[0] + v3[1] * v3[1] + v3[2] * v3[2]);
        for (double& v : v3) v /= s;
        double v1[3] = {J[0][0], J[0][1], J[0][2]};
        s = std::sqrt(v1[0] * v1[0] + v1[1] * v1[1] + v1[2] * v1[2]);
        for (double& v : v1) v /= s;
        double v2[3] = {v3[1] * v1[2] - v3[2] * v1[1], -v3[0] * v1[2] + v3[2] * v1[0], v3[0] * v1[1] - v3[1] * v1[0]};
        s = std::sqrt(v2[0] * v2[0] + v2[1] * v2[1] + v2[2] * v2[2]);
        for (double& v : v2) v /= s;
        const double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                           J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                           J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
        double iJ[3][3];
        iJ[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) / det;
        iJ[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) / det;
        iJ[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) / det;
        iJ[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) / det;
        iJ[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) / det;
        iJ[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) / det;
        iJ[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) / det;
        iJ[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) / det;
        iJ[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / det;
        const double wq = t.w[kx] * t.w[ky] * t.wz[kz] * det;
        // support functions and their B / Bfi columns
        S.clear();
        Bc.clear();
        Bf.clear();
        for (int j3 = 0; j3 < nz; ++j3)
          for (int j2 = 0; j2 < n; ++j2)
            for (int j1 = 0; j1 < n; ++j1) {
              const double dx = (j2 == ky && j3 == kz) ? t.D[kx * n + j1] : 0.0;
              const double dy = (j1 == kx && j3 == kz) ? t.D[ky * n + j2] : 0.0;
              const double dz = (j1 == kx && j2 == ky) ? t.Dz[kz * nz + j3] : 0.0;
              if (dx == 0.0 && dy == 0.0 && dz == 0.0) continue;
              // Nx = dN/dx etc. (d xi_r / d x_c = iJ[c][r])
              const double N[3] = {dx * iJ[0][0] + dy * iJ[0][1] + dz * iJ[0][2],
                                   dx * iJ[1][0] + dy * iJ[1][1] + dz * iJ[1][2],
                                   dx * iJ[2][0] + dy * iJ[2][1] + dz * iJ[2][2]};
              const double g1 = N[0] * v1[0] + N[1] * v1[1] + N[2] * v1[2];
              const double g2 = N[0] * v2[0] + N[1] * v2[1] + N[2] * v2[2];
              const double g3 = N[0] * v3[0] + N[1] * v3[1] + N[2] * v3[2];
              S.push_back((j3 * n + j2) * n + j1);
              double B[6][3];
              if (frame) {
                for (int c = 0; c < 3; ++c) {
                  B[0][c] = g1 * v1[c];
                  B[1][c] = g2 * v2[c];
                  B[2][c] = g3 * v3[c];
                  B[3][c] = g1 * v2[c] + g2 * v1[c];
                  B[4][c] = g2 * v3[c] + g3 * v2[c];
                  B[5][c] = g1 * v3[c] + g3 * v1[c];
                }
              } else {
                const double b[6][3] = {{N[0], 0, 0}, {0, N[1], 0}, {0, 0, N[2]},
                                        {N[1], N[0], 0}, {0, N[2], N[1]}, {N[2], 0, N[0]}};
                std::memcpy(B, b, sizeof(B));
              }
              for (int r = 0; r < 6; ++r)
                for (int c = 0; c < 3; ++c) Bc.push_back(B[r][c]);
              Bf.push_back(g1);
              Bf.push_back(g2);
              Bf.push_back(g3);
            }
        const int ns = (int)S.size();
        // E[a] = e^T Bfi[:, a]  (6-vector), G[a] = g Bfi[:, a]
        std::vector<double> EB((size_t)ns * 6), GB((size_t)ns * 3);
        for (int a = 0; a < ns; ++a) {
          const double* bf = &Bf[(size_t)a * 3];
          for (int r = 0; r < 6; ++r) EB[(size_t)a * 6 + r] = e[0 * 6 + r] * bf[0] + e[1 * 6 + r] * bf[1] + e[2 * 6 + r] * bf[2];
          for (int r = 0; r < 3; ++r) GB[(size_t)a * 3 + r] = g[r * 3 + 0] * bf[0] + g[r * 3 + 1] * bf[1] + g[r * 3 + 2] * bf[2];
        }
        for (int a = 0; a < ns; ++a) {
          const double* B = &Bc[(size_t)a * 18];
          for (int b = 0; b < ns; ++b) {
            const double* E = &EB[(size_t)b * 6];
            for (int c = 0; c < 3; ++c) {
              double v = 0.0;
              for (int r = 0; r < 6; ++r) v += B[r * 3 + c] * E[r];
              kuf[((size_t)3 * S[a] + c) * np + S[b]] += v * wq;
            }
            const double* bf = &Bf[(size_t)a * 3];
            const double* G = &GB[(size_t)b * 3];
            kff[(size_t)S[a] * np + S[b]] += (bf[0] * G[0] + bf[1] * G[1] + bf[2] * G[2]) * wq;
          }
        }
      }
}

// Assembled electrical matrices of one transducer over its node set (compact indices).
struct PztSystem {
  std::vector<int32_t> ids;         // compact -> caller's node
  std::map<int32_t, int> at;        // caller's node -> compact
  std::vector<int> ebc0, ebci, ebcn;  // compact indices, in ascending caller's node id
  std::vector<double> Kfu;          // [nn x 3nn]
  std::vector<double> Kff;          // [nn x nn]
  int nn = 0;
};

int assemble(int n_xy, int n_z, int64_t n_pel, const int32_t* pn, int64_t n_nodes, const double* coords,
             const double* e, const double* g, int frame, PztSystem& P) {
  if (n_xy < 2 || n_z < 2 || n_pel <= 0 || !pn || !coords || !e || !g) return SEM_E_ARG;
  const int np = n_xy * n_xy * n_z, nf = n_xy * n_xy;
  std::vector<int32_t> all(pn, pn + n_pel * np);
  for (int32_t v : all)
    if (v < 0 || v >= n_nodes) return SEM_E_ARG;
  std::vector<int32_t> u(all);
  std::sort(u.begin(), u.end());
  u.erase(std::unique(u.begin(), u.end()), u.end());
  if (u.size() > 20000) return SEM_E_UNSUPPORTED;  // dense electrical solve
  P.ids = u;
  P.nn = (int)u.size();
  for (int i = 0; i < P.nn; ++i) P.at[u[i]] = i;
  // electrode sets (main_spec3Dcuda_global_..._mat.m:200-223): bottom face, interior
  // levels, top face of every element; unique -> ascending node id
  std::vector<int32_t> s0, si, sn;
  for (int64_t el = 0; el < n_pel; ++el)
    for (int p = 0; p < np; ++p) {
      const int32_t nd = pn[el * np + p];
      (p < nf ? s0 : (p >= np - nf ? sn : si)).push_back(nd);
    }
  for (auto* s : {&s0, &si, &sn}) {
    std::sort(s->begin(), s->end());
    s->erase(std::unique(s->begin(), s->end()), s->end());
  }
  for (int32_t v : s0) P.ebc0.push_back(P.at[v]);
  for (int32_t v : si) P.ebci.push_back(P.at[v]);
  for (int32_t v : sn) P.ebcn.push_back(P.at[v]);
  const int nn = P.nn;
  P.Kfu.assign((size_t)nn * 3 * nn, 0.0);
  P.Kff.assign((size_t)nn * nn, 0.0);
  Gll t(n_xy, n_z);
  std::vector<double> X((size_t)np * 3), kuf, kff;
  std::vector<int> loc(np);
  for (int64_t el = 0; el < n_pel; ++el) {
    for (int p = 0; p < np; ++p) {
      const int32_t nd = pn[el * np + p];
      loc[p] = P.at[nd];
      for (int c = 0; c < 3; ++c) X[(size_t)p * 3 + c] = coords[(int64_t)nd * 3 + c];
    }
    pzt_element(t, X.data(), e, g, frame, kuf, kff);
    for (int a = 0; a < np; ++a)
      for (int b = 0; b < np; ++b) {
        P.Kff[(size_t)loc[a] * nn + loc[b]] += kff[(size_t)a * np + b];
        for (int c = 0; c < 3; ++c)  // Kfiu = kufi^T
          P.Kfu[(size_t)loc[b] * 3 * nn + 3 * loc[a] + c] += kuf[((size_t)3 * a + c) * np + b];
      }
  }
  return SEM_OK;
}

// FaL = Kfiu^T FiA with FiA = V0 on the top electrode, the induced potential inside and 0 on
// the bottom electrode (main_spec3Dcuda_global_..._mat.m:244-255); per compact node.
int actuator_forces(const PztSystem& P, double V0, std::vector<double>& fa) {
  const int nI = (int)P.ebci.size(), nn = P.nn;
  std::vector<double> fi(nn, 0.0);
  for (int c : P.ebcn) fi[c] = V0;
  if (nI > 0) {
    std::vector<double> A((size_t)nI * nI), rhs(nI, 0.0);
    for (int r = 0; r < nI; ++r) {
      for (int c = 0; c < nI; ++c) A[(size_t)r * nI + c] = P.Kff[(size_t)P.ebci[r] * nn + P.ebci[c]];
      double s = 0.0;
      for (int c : P.ebcn) s += P.Kff[(size_t)P.ebci[r] * nn + c] * V0;
      rhs[r] = -s;
    }
    if (!lu_solve(nI, A, 1, rhs)) return SEM_E_ARG;
    for (int r = 0; r < nI; ++r) fi[P.ebci[r]] = rhs[r];
  }
  fa.assign((size_t)3 * nn, 0.0);
  for (int a = 0; a < nn; ++a)
    if (fi[a] != 0.0)
      for (int j = 0; j < 3 * nn; ++j) fa[j] += P.Kfu[(size_t)a * 3 * nn + j] * fi[a];
  return SEM_OK;
}

}  // namespace

extern "C" {

int sem_pzt_const(int variant, const double* S, const double* d, const double* epsT, double theta_deg, double* Q,
                  double* e, double* g, double* e31_e32_g33) {
  if (!S || !d || !epsT || !Q || !e || !g) return SEM_E_ARG;
  // column order of the coupling matrix: (ex ey ez gxz gyz gxy) -> (ex ey ez gxy gxz gyz)
  const int perm[6] = {0, 1, 2, 5, 4, 3};
  double Sinv[36];
  if (variant == 1) {
    // solid3D/pzt_const.m:1-25: ep = dp S^-1, g = epsT - ep dp^T, rows and columns of S
    // rearranged before the inversion
    if (!inv6(S, Sinv)) return SEM_E_ARG;
    double ep[18];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 6; ++j) {
        double s = 0.0;
        for (int k = 0; k < 6; ++k) s += d[i * 6 + k] * Sinv[k * 6 + j];
        ep[i * 6 + j] = s;
      }
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        double s = 0.0;
        for (int k = 0; k < 6; ++k) s += ep[i * 6 + k] * d[j * 6 + k];
        g[i * 3 + j] = epsT[i * 3 + j] - s;
      }
    double Sr[36];
    for (int i = 0; i < 6; ++i)
      for (int j = 0; j < 6; ++j) Sr[i * 6 + j] = S[perm[i] * 6 + perm[j]];
    if (!inv6(Sr, Q)) return SEM_E_ARG;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 6; ++j) e[i * 6 + j] = ep[i * 6 + perm[j]];
    if (e31_e32_g33) {
      e31_e32_g33[0] = ep[12];
      e31_e32_g33[1] = ep[13];
      e31_e32_g33[2] = g[8];
    }
    return SEM_OK;
  }
  if (variant != 0) return SEM_E_ARG;
  // flat_shell_engine/pzt_const.m:1-37: charge constants rotated by theta (Tinv, eq. 2.31 of
  // Sierakowski), ep = dp_^T S^-1, g = epsT, Q = S^-1 in the original arrangement
  const double th = theta_deg * M_PI / 180.0, m = std::cos(th), n = std::sin(th);
  const double T[36] = {m * m, n * n, 0, 0, 0, -2 * m * n, n * n, m * m, 0, 0, 0, 2 * m * n,
                        0, 0, 1, 0, 0, 0, 0, 0, 0, m, n, 0,
                        0, 0, 0, -n, m, 0, m * n, -m * n, 0, 0, 0, m * m - n * n};
  double dr[18];  // dp_^T [3 x 6] = (T dp^T)^T
  for (int i = 0; i < 3; ++i)
    for (int r = 0; r < 6; ++r) {
      double s = 0.0;
      for (int k = 0; k < 6; ++k) s += T[r * 6 + k] * d[i * 6 + k];
      dr[i * 6 + r] = s;
    }
  if (!inv6(S, Sinv)) return SEM_E_ARG;
  double ep[18];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 6; ++j) {
      double s = 0.0;
      for (int k = 0; k < 6; ++k) s += dr[i * 6 + k] * Sinv[k * 6 + j];
      ep[i * 6 + j] = s;
    }
  std::copy(epsT, epsT + 9, g);
  std::copy(Sinv, Sinv + 36, Q);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 6; ++j) e[i * 6 + j] = ep[i * 6 + perm[j]];
  if (e31_e32_g33) {
    e31_e32_g33[0] = ep[12];
    e31_e32_g33[1] = ep[13];
    e31_e32_g33[2] = epsT[8];
  }
  return SEM_OK;
}

int sem_pzt_actuator(int n_xy, int n_z, int64_t n_pel, const int32_t* pnodes, int64_t n_nodes, const double* coords,
                     const double* e, const double* g, int frame, double V0, double* fa) {
  if (!fa) return SEM_E_ARG;
  PztSystem P;
  int rc = assemble(n_xy, n_z, n_pel, pnodes, n_nodes, coords, e, g, frame, P);
  if (rc != SEM_OK) return rc;
  std::vector<double> f;
  rc = actuator_forces(P, V0, f);
  if (rc != SEM_OK) return rc;
  std::fill(fa, fa + 3 * n_nodes, 0.0);
  for (int a = 0; a < P.nn; ++a)
    for (int c = 0; c < 3; ++c) fa[(int64_t)P.ids[a] * 3 + c] = f[(size_t)3 * a + c];
  return SEM_OK;
}

int sem_pzt_sensor(int n_xy, int n_z, int64_t n_pel, const int32_t* pnodes, int64_t n_nodes, const double* coords,
                   const double* e, const double* g, int frame, double* w) {
  if (!w) return SEM_E_ARG;
  PztSystem P;
  int rc = assemble(n_xy, n_z, n_pel, pnodes, n_nodes, coords, e, g, frame, P);
  if (rc != SEM_OK) return rc;
  // open circuit (main_spec3Dcuda_global_..._mat.m:356-361, 1062-1066): potentials of
  // [interior; top] nodes solve KfifiO fio = -KfiuO u, the voltage is the mean of the last
  // n_xy^2 entries of fio.  w = -KfiuO^T KfifiO^-T s with s that mean's selector.
  std::vector<int> O(P.ebci);
  O.insert(O.end(), P.ebcn.begin(), P.ebcn.end());
  const int nO = (int)O.size(), nn = P.nn, nf = n_xy * n_xy;
  if (nO < nf) return SEM_E_ARG;
  std::vector<double> A((size_t)nO * nO), s(nO, 0.0);
  for (int r = 0; r < nO; ++r)
    for (int c = 0; c < nO; ++c) A[(size_t)c * nO + r] = P.Kff[(size_t)O[r] * nn + O[c]];  // transpose
  for (int r = nO - nf; r < nO; ++r) s[r] = 1.0 / nf;
  if (!lu_solve(nO, A, 1, s)) return SEM_E_ARG;
  std::fill(w, w + 3 * n_nodes, 0.0);
  for (int r = 0; r < nO; ++r) {
    if (s[r] == 0.0) continue;
    const double* row = &P.Kfu[(size_t)O[r] * 3 * nn];
    for (int a = 0; a < nn; ++a)
      for (int c = 0; c < 3; ++c) w[(int64_t)P.ids[a] * 3 + c] -= s[r] * row[3 * a + c];
  }
  return SEM_OK;
}

int sem_pzt_shell_actuator(int n_xy, int n_z, int64_t n_act, const int32_t* act_nodes, int64_t n_nodes,
                           const double* coords, double t_pzt, double h, const double* e, const double* g, double V0,
                           double* fa) {
  // main_flat_shell_multi_pzt_c.m:158-235 with connect_pzt3D_to_plate_fun.m:1-44: a 3D
  // transducer of n_z GLL levels over the actuator's shell elements (level k at
  // z + (xi_k + 1)/2 t_pzt + h/2), its induced potential and forces; the in-plane forces of
  // the bottom-face boundary nodes act on the shell nodes, with moments (h/2 + t_pzt/2) F.
  if (n_xy < 2 || n_z < 2 || n_act <= 0 || !act_nodes || !coords || !e || !g || !fa) return SEM_E_ARG;
  const int nf = n_xy * n_xy, np = nf * n_z;
  std::vector<int32_t> bot(act_nodes, act_nodes + n_act * nf);
  for (int32_t v : bot)
    if (v < 0 || v >= n_nodes) return SEM_E_ARG;
  std::vector<int32_t> ub(bot);
  std::sort(ub.begin(), ub.end());
  ub.erase(std::unique(ub.begin(), ub.end()), ub.end());
  std::map<int32_t, int> rank;
  for (size_t i = 0; i < ub.size(); ++i) rank[ub[i]] = (int)i;
  // virtual nodes: level 0 = shell node ids, level k = n_nodes + (k-1)*|ub| + rank
  const int64_t nvirt = n_nodes + (int64_t)(n_z - 1) * ub.size();
  std::vector<double> xz, wz;
  gll(n_z, xz, wz);
  std::vector<double> vc((size_t)nvirt * 3, 0.0);
  std::vector<int32_t> pn((size_t)n_act * np);
  for (int64_t el = 0; el < n_act; ++el)
    for (int k = 0; k < n_z; ++k)
      for (int p = 0; p < nf; ++p) {
        const int32_t nd = act_nodes[el * nf + p];
        const int64_t id = k == 0 ? nd : n_nodes + (int64_t)(k - 1) * ub.size() + rank[nd];
        pn[(size_t)el * np + k * nf + p] = (int32_t)id;
        vc[(size_t)id * 3 + 0] = coords[(int64_t)nd * 3 + 0];
        vc[(size_t)id * 3 + 1] = coords[(int64_t)nd * 3 + 1];
        vc[(size_t)id * 3 + 2] = coords[(int64_t)nd * 3 + 2] + (xz[k] + 1.0) / 2.0 * t_pzt + h / 2.0;
      }
  if (nvirt > INT32_MAX) return SEM_E_ARG;
  PztSystem P;
  int rc = assemble(n_xy, n_z, n_act, pn.data(), nvirt, vc.data(), e, g, 0, P);
  if (rc != SEM_OK) return rc;
  std::vector<double> f;
  rc = actuator_forces(P, V0, f);
  if (rc != SEM_OK) return rc;
  // boundary nodes of the footprint (find_boundary_nodes.m:24-170): a corner node shared
  // with at most one other element's corners, an edge node on no other element's edges
  const int n = n_xy;
  auto is_corner = [&](int p) { return p == 0 || p == n - 1 || p == nf - n || p == nf - 1; };
  auto is_edge = [&](int p) {
    const int i = p % n, j = p / n;
    return !is_corner(p) && (i == 0 || i == n - 1 || j == 0 || j == n - 1);
  };
  std::vector<int32_t> outer;
  for (int64_t el = 0; el < n_act; ++el)
    for (int p = 0; p < nf; ++p) {
      if (!is_corner(p) && !is_edge(p)) continue;
      const int32_t nd = act_nodes[el * nf + p];
      int nb = 0;
      for (int64_t o = 0; o < n_act; ++o) {
        if (o == el) continue;
        for (int q = 0; q < nf; ++q)
          if (act_nodes[o * nf + q] == nd && (is_corner(p) ? is_corner(q) : is_edge(q))) ++nb;
      }
      if (is_corner(p) ? nb <= 1 : nb == 0) outer.push_back(nd);
    }
  std::sort(outer.begin(), outer.end());
  outer.erase(std::unique(outer.begin(), outer.end()), outer.end());
  std::fill(fa, fa + 5 * n_nodes, 0.0);
  const double arm = h / 2.0 + t_pzt / 2.0;
  for (int32_t nd : outer) {
    const int a = P.at.at(nd);
    const double fx = f[(size_t)3 * a], fy = f[(size_t)3 * a + 1];
    fa[(int64_t)nd * 5 + 0] = fx;
    fa[(int64_t)nd * 5 + 2] = fy;
    fa[(int64_t)nd * 5 + 1] = fx * arm;
    fa[(int64_t)nd * 5 + 3] = fy * arm;
  }
  return SEM_OK;
}

int sem_laminate_shell_add_pzt(const double* row, double h, double t_pzt, double rho_pzt, const double* Q,
                               double* out) {
  // composite_plate_c_delam_pzt.m:175-215: the transducer as an extra layer from h/2 to
  // h/2 + t_pzt/2 for stiffness (as coded), full thickness for m11; Q in the
  // (ex ey ez gxz gyz gxy) arrangement of the flat-shell pzt_const
  if (!row || !Q || !out) return SEM_E_ARG;
  std::copy(row, row + 23, out);
  const double h2 = h / 2 + t_pzt / 2, h1 = h / 2;
  const double sh = 5.0 / 4.0 * (h2 - h1 - 4.0 / 3.0 * (h2 * h2 * h2 - h1 * h1 * h1) / ((h + t_pzt) * (h + t_pzt)));
  const double A = h2 - h1, B = (h2 * h2 - h1 * h1) / 2, D = (h2 * h2 * h2 - h1 * h1 * h1) / 3;
  auto q = [&](int i, int j) { return Q[(i - 1) * 6 + (j - 1)]; };
  // a11 a12 a16 a22 a26 a66 a44 a45 a55 b11 b12 b16 b22 b26 b66 d11 d12 d16 d22 d26 d66 m11 I11
  const int ij[6][2] = {{1, 1}, {1, 2}, {1, 6}, {2, 2}, {2, 6}, {6, 6}};
  for (int k = 0; k < 6; ++k) {
    const double v = q(ij[k][0], ij[k][1]);
    out[k] += v * A;
    out[9 + k] += v * B;
    out[15 + k] += v * D;
  }
  out[6] += q(4, 4) * sh;
  out[7] += q(4, 5) * sh;
  out[8] += q(5, 5) * sh;
  out[21] += rho_pzt * t_pzt;
  out[22] += rho_pzt * (h2 * h2 * h2 - h1 * h1 * h1) / 3;
  return SEM_OK;
}

}  // extern "C"
