// Patch kernels of the spectral-element time step for gfx950 (structure: sem_kernels.hpp).
//
// Element work maps one thread to one GLL point; a workgroup holds the P elements of one
// patch so that every derivative contraction reads its operand lines from LDS.  T is the
// element arithmetic type (double for the plain path, float for the split-precision
// path); patch-node gathers, assembly sums and the state are always double.
#include "sem_kernels.hpp"

namespace semk {

// ---------------------------------------------------------------------------
// shared helpers
// ---------------------------------------------------------------------------
template <int N>
struct ShellP {
  static constexpr int NP = N * N;
  static constexpr int P = 16;
  static constexpr int TH = ((P * NP + 63) / 64) * 64;
};

template <int N, int NZ>
struct SolidP {
  static constexpr int NP = N * N * NZ;
  static constexpr int P = NP <= 128 ? 8 : (NP <= 256 ? 4 : 2);
  static constexpr int TH = ((P * NP + 63) / 64) * 64;
};

int shell_patch_elems(int, bool) { return 16; }
int solid_patch_elems(int n, int nz, bool) {
  const int np = n * n * nz;
  return np <= 128 ? 8 : (np <= 256 ? 4 : 2);
}
bool shell_supported(int n) { return n >= 3 && n <= 8; }
bool solid_supported(int n, int nz) { return (n >= 3 && n <= 6) && (nz == 3 || nz == n); }

size_t patch_lds_bytes(int kind, int n, int nz, bool fp32, int max_pn) {
  const int nf = kind == 1 ? 5 : 3, fd = kind == 1 ? 2 : 3;
  const int np = kind == 1 ? n * n : n * n * nz;
  const int P = kind == 1 ? 16 : solid_patch_elems(n, nz, fp32);
  const size_t ts = fp32 ? 4 : 8;
  return 8 * (size_t)max_pn * nf + ts * ((size_t)P * nf * np + (size_t)P * fd * np + n * n + nz * nz + np) + 16;
}

// Central-difference update of one node (force_excit_c.m:687-693 rearranged as
// U^{n+1} = c1*minv*(-F) + c2*(2U^n - U^{n-1}) + c3*U^{n-1}, with per-field constants),
// plus frame output and divergence guard (force_excit_c.m:710-730).
template <int NF, bool AK>
__device__ __forceinline__ void finish_node(const UpdateArgs& ua, int64_t node, const double* F,
                                            const double* Un, int64_t fi, bool frame, bool& bad) {
  const int64_t d0 = node * NF;
  if (AK) {
#pragma unroll
    for (int f = 0; f < NF; ++f) ua.Unew[d0 + f] = F[f];
    return;
  }
  double* fr = frame ? ua.frames + fi * ua.dof : nullptr;
#pragma unroll
  for (int f = 0; f < NF; ++f) {
    const double u = Un[f], uo = ua.Uold[d0 + f];
    const double un = ua.c1[f] * ua.minv[d0 + f] * (-F[f]) + ua.c2[f] * (2.0 * u - uo) + ua.c3[f] * uo;
    ua.Unew[d0 + f] = un;
    if (frame) {
      double v;
      switch (ua.frame_q) {
        case 0: v = u; break;
        case 1: v = ua.a1 * (-uo + un); break;
        case 2: v = un; break;
        default: v = ua.a0 * (uo - 2.0 * u + un); break;
      }
      fr[d0 + f] = v;
      if (ua.guard_limit > 0.0 && (ua.guard_field < 0 || ua.guard_field == f)) {
        const double g = ua.guard_post ? un : u;
        if (g > ua.guard_limit) bad = true;
      }
      if (d0 + f == 2 && un != un) bad = true;
    }
  }
}

// Patch assembly: every patch node sums its occurrences from the element-local forces
// in LDS (basket order); owned nodes are updated, boundary nodes write a partial.
template <int NF, int NP, int TH, bool AK, typename T>
__device__ __forceinline__ void patch_assemble(const PatchArgs& a, const UpdateArgs& ua, const T* su,
                                               const double* Up, int b, int pn0, int pn) {
  const uint16_t* optr = a.occ_ptr + pn0 + b;
  const uint16_t* oc = a.occ + a.occ_base[b];
  const int nown = a.pn_owned[b];
  const int64_t nn = AK ? 0 : ua.ctr[0];
  const int64_t fi = AK ? 0 : ua.ctr[1];
  const bool frame = !AK && fi < ua.n_frames && ua.frame_nn[fi] == nn;
  bool bad = false;
  for (int ln = threadIdx.x; ln < pn; ln += TH) {
    double F[NF];
#pragma unroll
    for (int f = 0; f < NF; ++f) F[f] = 0.0;
    const int o1 = optr[ln + 1];
    for (int o = optr[ln]; o < o1; ++o) {
      const int q = oc[o];
      const int le = q / NP, p = q - le * NP;
      const T* src = su + le * NF * NP + p;
#pragma unroll
      for (int f = 0; f < NF; ++f) F[f] += (double)src[f * NP];
    }
    const int64_t nd = a.pn_id[pn0 + ln];
    if (ln < nown) {
      finish_node<NF, AK>(ua, nd, F, Up + ln * NF, fi, frame, bad);
    } else {
      double* dst = ua.partial + (int64_t)(a.slot_base[b] + ln - nown) * NF;
#pragma unroll
      for (int f = 0; f < NF; ++f) dst[f] = F[f];
    }
  }
  if (bad) atomicMax((unsigned long long*)&ua.ctr[2], 1ull);
}

// ---------------------------------------------------------------------------
// flat shell, 5 DOF/node (u, phix, v, phiy, w)
// coefficient row: a11 a12 a16 a22 a26 a66 a44 a45 a55 b11 b12 b16 b22 b26 b66
//                  d11 d12 d16 d22 d26 d66 m11 I11
// ---------------------------------------------------------------------------
template <int N, typename T, bool AFF, bool AK>
__global__ __launch_bounds__(ShellP<N>::TH) void shell_patch_kernel(const double* __restrict__ gtab,
                                                                   PatchArgs a, UpdateArgs ua) {
  constexpr int NP = N * N, NF = 5, TH = ShellP<N>::TH;
  extern __shared__ double smem[];
  double* Up = smem;
  T* su = (T*)(Up + a.max_pn * NF);
  T* fl = su + ShellP<N>::P * NF * NP;
  T* sD = fl + ShellP<N>::P * 2 * NP;
  T* sW = sD + N * N;
  const int tid = threadIdx.x, b = blockIdx.x;
  const int pn0 = a.pn_base[b], pn = a.pn_base[b + 1] - pn0;
  const int e0 = a.el0[b], ne = a.el0[b + 1] - e0;
  for (int ln = tid; ln < pn; ln += TH) {
    const int64_t nd = a.pn_id[pn0 + ln];
#pragma unroll
    for (int f = 0; f < NF; ++f) Up[ln * NF + f] = ua.U[nd * NF + f];
  }
  if (tid < N * N) sD[tid] = (T)gtab[tid];
  if (tid < NP) sW[tid] = (T)(gtab[N * N + (tid % N)] * gtab[N * N + tid / N]);
  __syncthreads();
  const int le = tid / NP, p = tid - le * NP;
  const int i = p % N, j = p / N;
  const bool valid = le < ne;
  const int64_t e = e0 + le;
  T* sue = su + le * NF * NP;
  T* fle = fl + le * 2 * NP;
  if (valid) {
    const int ln = a.lmap[e * NP + p];
#pragma unroll
    for (int f = 0; f < NF; ++f) sue[f * NP + p] = (T)Up[ln * NF + f];
  }
  __syncthreads();
  T G[5][2];
  T dmx = 0, dmy = 0;
  if (valid) {
    T dxi[5], deta[5];
#pragma unroll
    for (int f = 0; f < 5; ++f) {
      T s1 = 0, s2 = 0;
#pragma unroll
      for (int k = 0; k < N; ++k) {
        s1 += sD[i * N + k] * sue[f * NP + j * N + k];
        s2 += sD[j * N + k] * sue[f * NP + k * N + i];
      }
      dxi[f] = s1;
      deta[f] = s2;
    }
    const T phx = sue[1 * NP + p], phy = sue[3 * NP + p];
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
    G[0][0] = Sxx * Wxx + Sxy * Wxy;   G[0][1] = Sxx * Wyx + Sxy * Wyy;    // fu
    G[1][0] = Sxxz * Wxx + Sxyz * Wxy; G[1][1] = Sxxz * Wyx + Sxyz * Wyy;  // mx
    G[2][0] = Sxy * Wxx + Syy * Wxy;   G[2][1] = Sxy * Wyx + Syy * Wyy;    // fv
    G[3][0] = Sxyz * Wxx + Syyz * Wxy; G[3][1] = Sxyz * Wyx + Syyz * Wyy;  // my
    G[4][0] = Sxz * Wxx + Syz * Wxy;   G[4][1] = Sxz * Wyx + Syz * Wyy;    // fw
    dmx = Sxz * W;
    dmy = Syz * W;
  }
  T out[5];
#pragma unroll
  for (int f = 0; f < 5; ++f) {
    if (valid) {
      fle[p] = G[f][0];
      fle[NP + p] = G[f][1];
    }
    __syncthreads();
    T s = 0;
    if (valid) {
#pragma unroll
      for (int k = 0; k < N; ++k) {
        s += sD[k * N + i] * fle[j * N + k];
        s += sD[k * N + j] * fle[NP + k * N + i];
      }
    }
    out[f] = s;
    __syncthreads();
  }
  if (valid) {
    out[1] += dmx;
    out[3] += dmy;
    if (!AK && a.excite && a.ex_ptr) {
      const int b0 = a.ex_ptr[e], b1 = a.ex_ptr[e + 1];
      if (b0 < b1) {
        const double stv = a.st[ua.ctr[0] - 1];
        for (int k = b0; k < b1; ++k) {
          const int pf = a.ex_pf[k];
          if (pf / NF == p) out[pf % NF] -= (T)(a.ex_amp[k] * stv);
        }
      }
    }
#pragma unroll
    for (int f = 0; f < NF; ++f) sue[f * NP + p] = out[f];
  }
  __syncthreads();
  patch_assemble<NF, NP, TH, AK, T>(a, ua, su, Up, b, pn0, pn);
}

// ---------------------------------------------------------------------------
// hexahedral solid, 3 DOF/node (u, v, w)
// coefficient row: D11 D12 D13 D14 D22 D23 D24 D33 D34 D44 D55 D56 D66 rho
// ---------------------------------------------------------------------------
template <int N, int NZ, typename T, bool AFF, bool AK>
__global__ __launch_bounds__((SolidP<N, NZ>::TH)) void solid_patch_kernel(const double* __restrict__ gtab,
                                                                         PatchArgs a, UpdateArgs ua) {
  constexpr int NP = N * N * NZ, NF = 3, TH = SolidP<N, NZ>::TH, P = SolidP<N, NZ>::P;
  extern __shared__ double smem[];
  double* Up = smem;
  T* su = (T*)(Up + a.max_pn * NF);
  T* fl = su + P * NF * NP;
  T* sD = fl + P * 3 * NP;
  T* sDz = sD + N * N;
  T* sW = sDz + NZ * NZ;
  const int tid = threadIdx.x, b = blockIdx.x;
  const int pn0 = a.pn_base[b], pn = a.pn_base[b + 1] - pn0;
  const int e0 = a.el0[b], ne = a.el0[b + 1] - e0;
  for (int ln = tid; ln < pn; ln += TH) {
    const int64_t nd = a.pn_id[pn0 + ln];
#pragma unroll
    for (int f = 0; f < NF; ++f) Up[ln * NF + f] = ua.U[nd * NF + f];
  }
  if (tid < N * N) sD[tid] = (T)gtab[tid];
  if (tid < NZ * NZ) sDz[tid] = (T)gtab[N * N + tid];
  if (tid < NP) {
    const double* w = gtab + N * N + NZ * NZ;
    const double* wz = w + N;
    const int ii = tid % N, jj = (tid / N) % N, kk = tid / (N * N);
    sW[tid] = (T)(w[ii] * w[jj] * wz[kk]);
  }
  __syncthreads();
  const int le = tid / NP, p = tid - le * NP;
  const int i = p % N, j = (p / N) % N, k = p / (N * N);
  const bool valid = le < ne;
  const int64_t e = e0 + le;
  T* sue = su + le * NF * NP;
  T* fle = fl + le * 3 * NP;
  if (valid) {
    const int ln = a.lmap[e * NP + p];
#pragma unroll
    for (int f = 0; f < NF; ++f) sue[f * NP + p] = (T)Up[ln * NF + f];
  }
  __syncthreads();
  T sig[3][3], iJ[3][3], W = 0;
  if (valid) {
    T dr[3][3];  // dr[f][d] = d u_f / d xi_d
#pragma unroll
    for (int f = 0; f < 3; ++f) {
      T s1 = 0, s2 = 0, s3 = 0;
#pragma unroll
      for (int m = 0; m < N; ++m) {
        s1 += sD[i * N + m] * sue[f * NP + (k * N + j) * N + m];
        s2 += sD[j * N + m] * sue[f * NP + (k * N + m) * N + i];
      }
#pragma unroll
      for (int m = 0; m < NZ; ++m) s3 += sDz[k * NZ + m] * sue[f * NP + (m * N + j) * N + i];
      dr[f][0] = s1;
      dr[f][1] = s2;
      dr[f][2] = s3;
    }
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
    sig[0][0] = Sxx; sig[0][1] = Sxy; sig[0][2] = Sxz;
    sig[1][0] = Sxy; sig[1][1] = Syy; sig[1][2] = Syz;
    sig[2][0] = Sxz; sig[2][1] = Syz; sig[2][2] = Szz;
  }
  T out[3];
#pragma unroll
  for (int f = 0; f < 3; ++f) {
    // flux G[d] = W * sum_c sig[f][c] iJ[d][c]  (main_spec3Dcuda_global_..._mat.m:944-954)
    if (valid) {
#pragma unroll
      for (int d = 0; d < 3; ++d)
        fle[d * NP + p] = W * (sig[f][0] * iJ[d][0] + sig[f][1] * iJ[d][1] + sig[f][2] * iJ[d][2]);
    }
    __syncthreads();
    T s = 0;
    if (valid) {
#pragma unroll
      for (int m = 0; m < N; ++m) {
        s += sD[m * N + i] * fle[(k * N + j) * N + m];
        s += sD[m * N + j] * fle[NP + (k * N + m) * N + i];
      }
#pragma unroll
      for (int m = 0; m < NZ; ++m) s += sDz[m * NZ + k] * fle[2 * NP + (m * N + j) * N + i];
    }
    out[f] = s;
    __syncthreads();
  }
  if (valid) {
    if (!AK && a.excite && a.ex_ptr) {
      const int b0 = a.ex_ptr[e], b1 = a.ex_ptr[e + 1];
      if (b0 < b1) {
        const double stv = a.st[ua.ctr[0] - 1];
        for (int q = b0; q < b1; ++q) {
          const int pf = a.ex_pf[q];
          if (pf / NF == p) out[pf % NF] -= (T)(a.ex_amp[q] * stv);
        }
      }
    }
#pragma unroll
    for (int f = 0; f < NF; ++f) sue[f * NP + p] = out[f];
  }
  __syncthreads();
  patch_assemble<NF, NP, TH, AK, T>(a, ua, su, Up, b, pn0, pn);
}

// ---------------------------------------------------------------------------
// patch-boundary nodes: partial sums + update; last block: outputs and counters
// ---------------------------------------------------------------------------
__device__ __forceinline__ double coherent_load(const double* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <int NF, bool AK>
__global__ __launch_bounds__(256) void shared_update_kernel(SharedArgs s, UpdateArgs ua) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nn = AK ? 0 : ua.ctr[0];
  const int64_t fi = AK ? 0 : ua.ctr[1];
  const bool frame = !AK && fi < ua.n_frames && ua.frame_nn[fi] == nn;
  bool bad = false;
  if (k < s.n_shared) {
    const int64_t nd = s.sh_id[k];
    double F[NF], Un[NF];
#pragma unroll
    for (int f = 0; f < NF; ++f) F[f] = 0.0;
    for (int o = s.sh_ptr[k]; o < s.sh_ptr[k + 1]; ++o) {
      const double* src = ua.partial + (int64_t)s.sh_slot[o] * NF;
#pragma unroll
      for (int f = 0; f < NF; ++f) F[f] += src[f];
    }
    if (!AK) {
#pragma unroll
      for (int f = 0; f < NF; ++f) Un[f] = ua.U[nd * NF + f];
    }
    finish_node<NF, AK>(ua, nd, F, Un, fi, frame, bad);
  }
  if (bad) atomicMax((unsigned long long*)&ua.ctr[2], 1ull);
  // last block to finish: every U^{n+1} of this step is written
  __shared__ int last;
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned long long prev = atomicAdd((unsigned long long*)&ua.ctr[4], 1ull);
    last = prev == (unsigned long long)gridDim.x - 1;
  }
  __syncthreads();
  if (!last) return;
  __threadfence();
  if (!AK && s.advance) {
    for (int jo = threadIdx.x; jo < s.n_out; jo += blockDim.x) {
      double acc = 0.0;
      for (int64_t q = s.out_ptr[jo]; q < s.out_ptr[jo + 1]; ++q) {
        const int64_t d = s.out_dof[q];
        double v;
        switch (s.out_q) {
          case 0: v = coherent_load(ua.U + d); break;
          case 1: v = ua.a1 * (-coherent_load(ua.Uold + d) + coherent_load(ua.Unew + d)); break;
          case 2: v = coherent_load(ua.Unew + d); break;
          default:
            v = ua.a0 * (coherent_load(ua.Uold + d) - 2.0 * coherent_load(ua.U + d) + coherent_load(ua.Unew + d));
            break;
        }
        acc += s.out_w[q] * v;
      }
      s.traces[jo * s.ld + (nn - s.nn_base)] = acc;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    ua.ctr[4] = 0;
    if (!AK && s.advance) {
      if (frame) ua.ctr[1] = fi + 1;
      if (__hip_atomic_load(&ua.ctr[2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0 && ua.ctr[3] == 0)
        ua.ctr[3] = nn;
      ua.ctr[0] = nn + 1;
    }
  }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
// Kernels that need more than 64 KB of dynamic LDS opt in once per size increase.
template <typename K>
static hipError_t opt_in(K kernel, size_t lds, size_t& granted) {
  if (lds <= 65536 || lds <= granted) return hipSuccess;
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e == hipSuccess) granted = lds;
  return e;
}

template <int N, typename T, bool AFF, bool AK>
static hipError_t shell_go(int64_t n_patch, const double* gtab, const PatchArgs& pa, const UpdateArgs& ua,
                           hipStream_t s) {
  const size_t lds = patch_lds_bytes(1, N, 1, sizeof(T) == 4, pa.max_pn);
  static size_t granted = 0;
  hipError_t e = opt_in(shell_patch_kernel<N, T, AFF, AK>, lds, granted);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((shell_patch_kernel<N, T, AFF, AK>), dim3((unsigned)n_patch), dim3(ShellP<N>::TH), lds, s,
                     gtab, pa, ua);
  return hipGetLastError();
}

template <int N, int NZ, typename T, bool AFF, bool AK>
static hipError_t solid_go(int64_t n_patch, const double* gtab, const PatchArgs& pa, const UpdateArgs& ua,
                           hipStream_t s) {
  const size_t lds = patch_lds_bytes(2, N, NZ, sizeof(T) == 4, pa.max_pn);
  static size_t granted = 0;
  hipError_t e = opt_in(solid_patch_kernel<N, NZ, T, AFF, AK>, lds, granted);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((solid_patch_kernel<N, NZ, T, AFF, AK>), dim3((unsigned)n_patch), dim3(SolidP<N, NZ>::TH),
                     lds, s, gtab, pa, ua);
  return hipGetLastError();
}

template <int N, int NZ, bool SHELL>
static hipError_t dispatch(bool fp32, bool aff, bool ak, int64_t np, const double* g, const PatchArgs& pa,
                           const UpdateArgs& ua, hipStream_t s) {
#define GO(T, A, K)                                           \
  do {                                                        \
    if constexpr (SHELL) return shell_go<N, T, A, K>(np, g, pa, ua, s); \
    else return solid_go<N, NZ, T, A, K>(np, g, pa, ua, s);  \
  } while (0)
  if (fp32) {
    if (aff) { if (ak) GO(float, true, true); else GO(float, true, false); }
    else { if (ak) GO(float, false, true); else GO(float, false, false); }
  } else {
    if (aff) { if (ak) GO(double, true, true); else GO(double, true, false); }
    else { if (ak) GO(double, false, true); else GO(double, false, false); }
  }
#undef GO
}

hipError_t launch_patch(int kind, int n, int nz, bool fp32, bool affine, bool applyK, int64_t n_patch,
                        const double* gtab, const PatchArgs& pa, const UpdateArgs& ua, hipStream_t s) {
  if (kind == 1) {
    switch (n) {
      case 3: return dispatch<3, 1, true>(fp32, affine, applyK, n_patch, gtab, pa, ua, s);
      case 4: return dispatch<4, 1, true>(fp32, affine, applyK, n_patch, gtab, pa, ua, s);
      case 5: return dispatch<5, 1, true>(fp32, affine, applyK, n_patch, gtab, pa, ua, s);
      case 6: return dispatch<6, 1, true>(fp32, affine, applyK, n_patch, gtab, pa, ua, s);
      case 7: return dispatch<7, 1, true>(fp32, affine, applyK, n_patch, gtab, pa, ua, s);
      case 8: return dispatch<8, 1, true>(fp32, affine, applyK, n_patch, gtab, pa, ua, s);
      default: return hipErrorInvalidValue;
    }
  }
#define SOLID_CASE(A, B) \
  if (n == A && nz == B) return dispatch<A, B, false>(fp32, affine, applyK, n_patch, gtab, pa, ua, s);
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

hipError_t launch_shared(int nf, bool applyK, const SharedArgs& sa, const UpdateArgs& ua, hipStream_t s) {
  const unsigned blocks = (unsigned)(sa.n_shared > 0 ? (sa.n_shared + 255) / 256 : 1);
#define SH(NFV, AK) hipLaunchKernelGGL((shared_update_kernel<NFV, AK>), dim3(blocks), dim3(256), 0, s, sa, ua)
  if (nf == 5) { if (applyK) SH(5, true); else SH(5, false); }
  else if (nf == 3) { if (applyK) SH(3, true); else SH(3, false); }
  else return hipErrorInvalidValue;
#undef SH
  return hipGetLastError();
}

}  // namespace semk
