// Element, node and post-step kernels for gfx950 (see sem_kernels.hpp for the step
// structure).  Element kernels map one thread to one GLL point; a block holds E
// whole elements so that every derivative contraction reads its operand lines from
// LDS.  T is the element arithmetic type (double for the plain path, float for the
// split-precision path); assembly sums and the state are always double.
#include "sem_kernels.hpp"

namespace semk {

// ---------------------------------------------------------------------------
// flat shell, 5 DOF/node (u, phix, v, phiy, w)
// ---------------------------------------------------------------------------
template <int N>
struct ShellCfg {
  static constexpr int NP = N * N;
  static constexpr int E = (576 / NP) > 0 ? (576 / NP) : 1;
  static constexpr int THREADS = E * NP;
};

// Shell coefficient row: a11 a12 a16 a22 a26 a66 a44 a45 a55 b11 b12 b16 b22 b26 b66
//                        d11 d12 d16 d22 d26 d66 m11 I11
template <int N, typename T, bool AFF>
__global__ __launch_bounds__(ShellCfg<N>::THREADS) void shell_forces_kernel(
    const double* __restrict__ gtab, ForceArgs a) {
  constexpr int NP = N * N, E = ShellCfg<N>::E;
  __shared__ T sbuf[E * 10 * NP];
  __shared__ T sD[N * N];
  __shared__ T sW[NP];
  const int tid = threadIdx.x;
  const int le = tid / NP, p = tid - le * NP;
  const int i = p % N, j = p / N;
  const int64_t e = (int64_t)blockIdx.x * E + le;
  const bool valid = e < a.n_el;
  if (tid < N * N) sD[tid] = (T)gtab[tid];
  if (tid < NP) sW[tid] = (T)(gtab[N * N + (tid % N)] * gtab[N * N + tid / N]);
  T* su = sbuf + le * 10 * NP;
  if (valid) {
    const int64_t nd = a.nodes[e * NP + p];
    const double* up = a.U + nd * 5;
#pragma unroll
    for (int f = 0; f < 5; ++f) su[f * NP + p] = (T)up[f];
  }
  __syncthreads();
  T dxi[5], deta[5], phx = 0, phy = 0;
  if (valid) {
#pragma unroll
    for (int f = 0; f < 5; ++f) {
      T s1 = 0, s2 = 0;
#pragma unroll
      for (int k = 0; k < N; ++k) {
        s1 += sD[i * N + k] * su[f * NP + j * N + k];
        s2 += sD[j * N + k] * su[f * NP + k * N + i];
      }
      dxi[f] = s1;
      deta[f] = s2;
    }
    phx = su[1 * NP + p];
    phy = su[3 * NP + p];
  }
  __syncthreads();
  T dmx = 0, dmy = 0;
  if (valid) {
    T xx, xy, yx, yy, W;  // xi_x, xi_y, eta_x, eta_y (reference invJ11, invJ12, invJ21, invJ22)
    if (AFF) {
      const double* g = a.geom + e * 5;
      xx = (T)g[0]; xy = (T)g[1]; yx = (T)g[2]; yy = (T)g[3];
      W = sW[p] * (T)g[4];
    } else {
      const double* g = a.geom + (e * NP + p) * 5;
      xx = (T)g[0]; xy = (T)g[1]; yx = (T)g[2]; yy = (T)g[3];
      W = (T)g[4];
    }
    const double* m = a.mat + (int64_t)a.elem_mat[e] * a.n_coef;
    const double sc = a.sscale ? a.sscale[e] : 1.0;
    const T a11 = (T)(m[0] * sc), a12 = (T)(m[1] * sc), a16 = (T)(m[2] * sc), a22 = (T)(m[3] * sc),
            a26 = (T)(m[4] * sc), a66 = (T)(m[5] * sc), a44 = (T)(m[6] * sc), a45 = (T)(m[7] * sc),
            a55 = (T)(m[8] * sc), b11 = (T)(m[9] * sc), b12 = (T)(m[10] * sc), b16 = (T)(m[11] * sc),
            b22 = (T)(m[12] * sc), b26 = (T)(m[13] * sc), b66 = (T)(m[14] * sc), d11 = (T)(m[15] * sc),
            d12 = (T)(m[16] * sc), d16 = (T)(m[17] * sc), d22 = (T)(m[18] * sc), d26 = (T)(m[19] * sc),
            d66 = (T)(m[20] * sc);
    // physical derivatives d/dx = d/dxi * xi_x + d/deta * eta_x (force_excit_c.m:615-622)
    const T u_x = dxi[0] * xx + deta[0] * yx, u_y = dxi[0] * xy + deta[0] * yy;
    const T fx_x = dxi[1] * xx + deta[1] * yx, fx_y = dxi[1] * xy + deta[1] * yy;
    const T v_x = dxi[2] * xx + deta[2] * yx, v_y = dxi[2] * xy + deta[2] * yy;
    const T fy_x = dxi[3] * xx + deta[3] * yx, fy_y = dxi[3] * xy + deta[3] * yy;
    const T w_x = dxi[4] * xx + deta[4] * yx, w_y = dxi[4] * xy + deta[4] * yy;
    const T ex = u_x, ey = v_y, gxy = u_y + v_x;
    const T kx = fx_x, ky = fy_y, kxy = fx_y + fy_x;
    const T gyz = phy + w_y, gxz = phx + w_x;  // '+' signs as coded in the reference
    const T Sxx = a11 * ex + a12 * ey + a16 * gxy + b11 * kx + b12 * ky + b16 * kxy;
    const T Syy = a12 * ex + a22 * ey + a26 * gxy + b12 * kx + b22 * ky + b26 * kxy;
    const T Sxy = a16 * ex + a26 * ey + a66 * gxy + b16 * kx + b26 * ky + b66 * kxy;
    const T Sxxz = b11 * ex + b12 * ey + b16 * gxy + d11 * kx + d12 * ky + d16 * kxy;
    const T Syyz = b12 * ex + b22 * ey + b26 * gxy + d12 * kx + d22 * ky + d26 * kxy;
    const T Sxyz = b16 * ex + b26 * ey + b66 * gxy + d16 * kx + d26 * ky + d66 * kxy;
    const T Syz = a44 * gyz + a45 * gxz;
    const T Sxz = a45 * gyz + a55 * gxz;
    // fluxes contracted by the transposed derivatives (force_excit_c.m:624-630)
    const T Wxx = W * xx, Wxy = W * xy, Wyx = W * yx, Wyy = W * yy;
    su[0 * NP + p] = Sxx * Wxx + Sxy * Wxy;   // fu, xi
    su[1 * NP + p] = Sxx * Wyx + Sxy * Wyy;   // fu, eta
    su[2 * NP + p] = Sxxz * Wxx + Sxyz * Wxy; // mx, xi
    su[3 * NP + p] = Sxxz * Wyx + Sxyz * Wyy; // mx, eta
    su[4 * NP + p] = Sxy * Wxx + Syy * Wxy;   // fv, xi
    su[5 * NP + p] = Sxy * Wyx + Syy * Wyy;   // fv, eta
    su[6 * NP + p] = Sxyz * Wxx + Syyz * Wxy; // my, xi
    su[7 * NP + p] = Sxyz * Wyx + Syyz * Wyy; // my, eta
    su[8 * NP + p] = Sxz * Wxx + Syz * Wxy;   // fw, xi
    su[9 * NP + p] = Sxz * Wyx + Syz * Wyy;   // fw, eta
    dmx = Sxz * W;
    dmy = Syz * W;
  }
  __syncthreads();
  if (valid) {
    T out[5];
#pragma unroll
    for (int f = 0; f < 5; ++f) {
      T s = 0;
#pragma unroll
      for (int k = 0; k < N; ++k) {
        s += sD[k * N + i] * su[(2 * f) * NP + j * N + k];
        s += sD[k * N + j] * su[(2 * f + 1) * NP + k * N + i];
      }
      out[f] = s;
    }
    out[1] += dmx;
    out[3] += dmy;
    if (a.excite && a.ex_ptr) {
      const int b0 = a.ex_ptr[e], b1 = a.ex_ptr[e + 1];
      if (b0 < b1) {
        const double stv = a.st[a.ctr[0] - 1];
        for (int k = b0; k < b1; ++k) {
          const int pf = a.ex_pf[k];
          if (pf / 5 == p) out[pf % 5] -= (T)(a.ex_amp[k] * stv);
        }
      }
    }
    T* fo = (T*)a.floc + (e * NP + p) * 5;
#pragma unroll
    for (int f = 0; f < 5; ++f) fo[f] = out[f];
  }
}

// ---------------------------------------------------------------------------
// hexahedral solid, 3 DOF/node (u, v, w)
// ---------------------------------------------------------------------------
template <int N, int NZ>
struct SolidCfg {
  static constexpr int NP = N * N * NZ;
  static constexpr int E = (512 / NP) > 0 ? (512 / NP) : 1;
  static constexpr int THREADS = E * NP;
};

// Solid coefficient row: D11 D12 D13 D14 D22 D23 D24 D33 D34 D44 D55 D56 D66 rho
template <int N, int NZ, typename T, bool AFF>
__global__ __launch_bounds__((SolidCfg<N, NZ>::THREADS)) void solid_forces_kernel(
    const double* __restrict__ gtab, ForceArgs a) {
  constexpr int NP = N * N * NZ, E = SolidCfg<N, NZ>::E;
  __shared__ T sbuf[E * 9 * NP];
  __shared__ T sD[N * N];
  __shared__ T sDz[NZ * NZ];
  __shared__ T sW[NP];
  const int tid = threadIdx.x;
  const int le = tid / NP, p = tid - le * NP;
  const int i = p % N, j = (p / N) % N, k = p / (N * N);
  const int64_t e = (int64_t)blockIdx.x * E + le;
  const bool valid = e < a.n_el;
  if (tid < N * N) sD[tid] = (T)gtab[tid];
  if (tid < NZ * NZ) sDz[tid] = (T)gtab[N * N + tid];
  if (tid < NP) {
    const double* w = gtab + N * N + NZ * NZ;
    const double* wz = w + N;
    const int ii = tid % N, jj = (tid / N) % N, kk = tid / (N * N);
    sW[tid] = (T)(w[ii] * w[jj] * wz[kk]);
  }
  T* su = sbuf + le * 9 * NP;
  if (valid) {
    const int64_t nd = a.nodes[e * NP + p];
    const double* up = a.U + nd * 3;
#pragma unroll
    for (int f = 0; f < 3; ++f) su[f * NP + p] = (T)up[f];
  }
  __syncthreads();
  T dr[3][3];  // dr[f][d] = d u_f / d xi_d
  if (valid) {
#pragma unroll
    for (int f = 0; f < 3; ++f) {
      T s1 = 0, s2 = 0, s3 = 0;
#pragma unroll
      for (int m = 0; m < N; ++m) {
        s1 += sD[i * N + m] * su[f * NP + (k * N + j) * N + m];
        s2 += sD[j * N + m] * su[f * NP + (k * N + m) * N + i];
      }
#pragma unroll
      for (int m = 0; m < NZ; ++m) s3 += sDz[k * NZ + m] * su[f * NP + (m * N + j) * N + i];
      dr[f][0] = s1;
      dr[f][1] = s2;
      dr[f][2] = s3;
    }
  }
  __syncthreads();
  if (valid) {
    T iJ[3][3], W;  // iJ[d][c] = d xi_d / d x_c (reference invJ(d+1,c+1))
    const double* g = AFF ? (a.geom + e * 10) : (a.geom + (e * NP + p) * 10);
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
      for (int c = 0; c < 3; ++c) iJ[d][c] = (T)g[3 * d + c];
    W = AFF ? sW[p] * (T)g[9] : (T)g[9];
    const double* m = a.mat + (int64_t)a.elem_mat[e] * a.n_coef;
    const double sc = a.sscale ? a.sscale[e] : 1.0;
    const T D11 = (T)(m[0] * sc), D12 = (T)(m[1] * sc), D13 = (T)(m[2] * sc), D14 = (T)(m[3] * sc),
            D22 = (T)(m[4] * sc), D23 = (T)(m[5] * sc), D24 = (T)(m[6] * sc), D33 = (T)(m[7] * sc),
            D34 = (T)(m[8] * sc), D44 = (T)(m[9] * sc), D55 = (T)(m[10] * sc), D56 = (T)(m[11] * sc),
            D66 = (T)(m[12] * sc);
    T gr[3][3];  // gr[f][c] = d u_f / d x_c (strains_spec_p.m:8-27)
#pragma unroll
    for (int f = 0; f < 3; ++f)
#pragma unroll
      for (int c = 0; c < 3; ++c) gr[f][c] = dr[f][0] * iJ[0][c] + dr[f][1] * iJ[1][c] + dr[f][2] * iJ[2][c];
    const T Exx = gr[0][0], Eyy = gr[1][1], Ezz = gr[2][2];
    const T Exy = gr[0][1] + gr[1][0], Eyz = gr[1][2] + gr[2][1], Exz = gr[0][2] + gr[2][0];
    // stresses_spec_p.m:6-11
    const T Sxx = Exx * D11 + Eyy * D12 + Ezz * D13 + Exy * D14;
    const T Syy = Exx * D12 + Eyy * D22 + Ezz * D23 + Exy * D24;
    const T Szz = Exx * D13 + Eyy * D23 + Ezz * D33 + Exy * D34;
    const T Sxy = Exx * D14 + Eyy * D24 + Ezz * D34 + Exy * D44;
    const T Syz = Eyz * D55 + Exz * D56;
    const T Sxz = Eyz * D56 + Exz * D66;
    const T sig[3][3] = {{Sxx, Sxy, Sxz}, {Sxy, Syy, Syz}, {Sxz, Syz, Szz}};
    // flux G[f][d] = W * sum_c sig[f][c] iJ[d][c]  (main_spec3Dcuda_global_..._mat.m:944-954)
#pragma unroll
    for (int f = 0; f < 3; ++f)
#pragma unroll
      for (int d = 0; d < 3; ++d)
        su[(3 * f + d) * NP + p] = W * (sig[f][0] * iJ[d][0] + sig[f][1] * iJ[d][1] + sig[f][2] * iJ[d][2]);
  }
  __syncthreads();
  if (valid) {
    T out[3];
#pragma unroll
    for (int f = 0; f < 3; ++f) {
      T s = 0;
#pragma unroll
      for (int m = 0; m < N; ++m) {
        s += sD[m * N + i] * su[(3 * f + 0) * NP + (k * N + j) * N + m];
        s += sD[m * N + j] * su[(3 * f + 1) * NP + (k * N + m) * N + i];
      }
#pragma unroll
      for (int m = 0; m < NZ; ++m) s += sDz[m * NZ + k] * su[(3 * f + 2) * NP + (m * N + j) * N + i];
      out[f] = s;
    }
    if (a.excite && a.ex_ptr) {
      const int b0 = a.ex_ptr[e], b1 = a.ex_ptr[e + 1];
      if (b0 < b1) {
        const double stv = a.st[a.ctr[0] - 1];
        for (int q = b0; q < b1; ++q) {
          const int pf = a.ex_pf[q];
          if (pf / 3 == p) out[pf % 3] -= (T)(a.ex_amp[q] * stv);
        }
      }
    }
    T* fo = (T*)a.floc + (e * NP + p) * 3;
#pragma unroll
    for (int f = 0; f < 3; ++f) fo[f] = out[f];
  }
}

// ---------------------------------------------------------------------------
// node kernel: deterministic assembly + central-difference update
// ---------------------------------------------------------------------------
template <int NF, typename T, bool APPLYK>
__global__ __launch_bounds__(256) void node_update_kernel(NodeArgs a) {
  const int64_t node = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (node >= a.n_nodes) return;
  double F[NF];
#pragma unroll
  for (int f = 0; f < NF; ++f) F[f] = 0.0;
  const int32_t o0 = a.rowptr[node], o1 = a.rowptr[node + 1];
  const T* fl = (const T*)a.floc;
  for (int32_t o = o0; o < o1; ++o) {
    const int64_t q = a.occ[o];
#pragma unroll
    for (int f = 0; f < NF; ++f) F[f] += (double)fl[q * NF + f];
  }
  const int64_t d0 = node * NF;
  if (APPLYK) {
#pragma unroll
    for (int f = 0; f < NF; ++f) a.Unew[d0 + f] = F[f];
    return;
  }
  const int64_t nn = a.ctr[0];
  const int64_t fi = a.ctr[1];
  const bool frame = fi < a.n_frames && a.frame_nn[fi] == nn;
  double* fr = frame ? a.frames + fi * (a.n_nodes * NF) : nullptr;
  bool bad = false;
#pragma unroll
  for (int f = 0; f < NF; ++f) {
    const double u = a.U[d0 + f], uo = a.Uold[d0 + f];
    const double un = a.c1[f] * a.minv[d0 + f] * (-F[f]) + a.c2[f] * (2.0 * u - uo) + a.c3[f] * uo;
    a.Unew[d0 + f] = un;
    if (frame) {
      double v;
      switch (a.frame_q) {
        case 0: v = u; break;
        case 1: v = a.a1 * (-uo + un); break;
        case 2: v = un; break;
        default: v = a.a0 * (uo - 2.0 * u + un); break;
      }
      fr[d0 + f] = v;
      if (a.guard_limit > 0.0 && (a.guard_field < 0 || a.guard_field == f)) {
        const double g = a.guard_post ? un : u;
        if (g > a.guard_limit) bad = true;
      }
      if (d0 + f == 2 && un != un) bad = true;
    }
  }
  if (bad) atomicMax((unsigned long long*)&a.ctr[2], 1ull);
}

// ---------------------------------------------------------------------------
// post-step: output functionals and step counters (one block)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void post_step_kernel(PostArgs a) {
  const int64_t nn = a.ctr[0];
  for (int jo = threadIdx.x; jo < a.n_out; jo += blockDim.x) {
    double s = 0.0;
    for (int64_t q = a.out_ptr[jo]; q < a.out_ptr[jo + 1]; ++q) {
      const int64_t d = a.out_dof[q];
      double v;
      switch (a.out_q) {
        case 0: v = a.U[d]; break;
        case 1: v = a.a1 * (-a.Uold[d] + a.Unew[d]); break;
        case 2: v = a.Unew[d]; break;
        default: v = a.a0 * (a.Uold[d] - 2.0 * a.U[d] + a.Unew[d]); break;
      }
      s += a.out_w[q] * v;
    }
    a.traces[jo * a.ld + (nn - a.nn_base)] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const int64_t fi = a.ctr[1];
    if (fi < a.n_frames && a.frame_nn[fi] == nn) a.ctr[1] = fi + 1;
    if (a.ctr[2] != 0 && a.ctr[3] == 0) a.ctr[3] = nn;
    a.ctr[0] = nn + 1;
  }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
template <int N, typename T, bool AFF>
static hipError_t shell_go(const double* gtab, const ForceArgs& a, hipStream_t s) {
  using C = ShellCfg<N>;
  const int64_t blocks = (a.n_el + C::E - 1) / C::E;
  hipLaunchKernelGGL((shell_forces_kernel<N, T, AFF>), dim3((unsigned)blocks), dim3(C::THREADS), 0, s,
                     gtab, a);
  return hipGetLastError();
}

template <int N>
static hipError_t shell_dispatch(bool fp32, bool aff, const double* gtab, const ForceArgs& a,
                                 hipStream_t s) {
  if (fp32) return aff ? shell_go<N, float, true>(gtab, a, s) : shell_go<N, float, false>(gtab, a, s);
  return aff ? shell_go<N, double, true>(gtab, a, s) : shell_go<N, double, false>(gtab, a, s);
}

bool shell_supported(int n) { return n >= 3 && n <= 8; }

hipError_t launch_shell_forces_tab(int n, bool fp32, bool affine, const double* gtab,
                                   const ForceArgs& a, hipStream_t s) {
  switch (n) {
    case 3: return shell_dispatch<3>(fp32, affine, gtab, a, s);
    case 4: return shell_dispatch<4>(fp32, affine, gtab, a, s);
    case 5: return shell_dispatch<5>(fp32, affine, gtab, a, s);
    case 6: return shell_dispatch<6>(fp32, affine, gtab, a, s);
    case 7: return shell_dispatch<7>(fp32, affine, gtab, a, s);
    case 8: return shell_dispatch<8>(fp32, affine, gtab, a, s);
    default: return hipErrorInvalidValue;
  }
}

template <int N, int NZ, typename T, bool AFF>
static hipError_t solid_go(const double* gtab, const ForceArgs& a, hipStream_t s) {
  using C = SolidCfg<N, NZ>;
  const int64_t blocks = (a.n_el + C::E - 1) / C::E;
  hipLaunchKernelGGL((solid_forces_kernel<N, NZ, T, AFF>), dim3((unsigned)blocks), dim3(C::THREADS), 0,
                     s, gtab, a);
  return hipGetLastError();
}

template <int N, int NZ>
static hipError_t solid_dispatch(bool fp32, bool aff, const double* gtab, const ForceArgs& a,
                                 hipStream_t s) {
  if (fp32) return aff ? solid_go<N, NZ, float, true>(gtab, a, s) : solid_go<N, NZ, float, false>(gtab, a, s);
  return aff ? solid_go<N, NZ, double, true>(gtab, a, s) : solid_go<N, NZ, double, false>(gtab, a, s);
}

bool solid_supported(int n, int nz) {
  return (n >= 3 && n <= 6) && (nz == 3 || nz == n);
}

hipError_t launch_solid_forces_tab(int n, int nz, bool fp32, bool affine, const double* gtab,
                                   const ForceArgs& a, hipStream_t s) {
#define SOLID_CASE(A, B) \
  if (n == A && nz == B) return solid_dispatch<A, B>(fp32, affine, gtab, a, s);
  SOLID_CASE(3, 3)
  SOLID_CASE(4, 4)
  SOLID_CASE(4, 3)
  SOLID_CASE(5, 5)
  SOLID_CASE(5, 3)
  SOLID_CASE(6, 6)
  SOLID_CASE(6, 3)
#undef SOLID_CASE
  return hipErrorInvalidValue;
}

hipError_t launch_node_update(int nf, bool fp32, bool applyK, const NodeArgs& a, hipStream_t s) {
  const unsigned blocks = (unsigned)((a.n_nodes + 255) / 256);
#define NODE_GO(NFV, TT, AK) \
  hipLaunchKernelGGL((node_update_kernel<NFV, TT, AK>), dim3(blocks), dim3(256), 0, s, a)
  if (nf == 5) {
    if (fp32) { if (applyK) NODE_GO(5, float, true); else NODE_GO(5, float, false); }
    else { if (applyK) NODE_GO(5, double, true); else NODE_GO(5, double, false); }
  } else if (nf == 3) {
    if (fp32) { if (applyK) NODE_GO(3, float, true); else NODE_GO(3, float, false); }
    else { if (applyK) NODE_GO(3, double, true); else NODE_GO(3, double, false); }
  } else {
    return hipErrorInvalidValue;
  }
#undef NODE_GO
  return hipGetLastError();
}

hipError_t launch_post_step(const PostArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(post_step_kernel, dim3(1), dim3(256), 0, s, a);
  return hipGetLastError();
}

}  // namespace semk
