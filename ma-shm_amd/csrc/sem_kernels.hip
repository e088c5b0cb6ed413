// Patch kernels of the spectral-element time step for gfx950 (structure: sem_kernels.hpp).
//
// Shells map one thread to one GLL point and hold the 16 elements of a 4x4 patch per
// workgroup.  Solids map one wavefront to one element (PPL points per lane), so that
// the element's geometry and material are wavefront-uniform and every exchange of the
// sum-factorised contractions stays inside the wavefront (no workgroup barrier); the
// workgroup holds the P elements of one patch.  T is the element arithmetic type
// (double for the plain path, float for the split-precision path); patch-node gathers,
// assembly sums and the state are always double.
#include "sem_device.hpp"

namespace semk {

int shell_patch_elems(int, bool) { return 16; }
int solid_patch_elems(int, int, bool) { return 4; }
bool shell_supported(int n) { return n >= 3 && n <= 8; }
bool solid_supported(int n, int nz) { return (n >= 3 && n <= 6) && (nz == 3 || nz == n); }

bool variant_accepts(int kind, int variant, int patch_elems) {
  if (variant == 0) return patch_elems == (kind == 1 ? 16 : 4);
  if (variant == 2) return kind == 1 && patch_elems == 16;
  if (variant == 1) return kind == 2 ? (patch_elems == 4 || patch_elems == 8) : (patch_elems == 8 || patch_elems == 16);
  if (variant == 3) return kind == 2 && (patch_elems == 4 || patch_elems == 8);
  return false;
}

// Dynamic LDS of a patch kernel; the static LDS of every patch kernel (the update
// constants, 3 x 5 doubles at most) is kStaticLds beside it.
constexpr size_t kStaticLds = 128;

size_t patch_lds_bytes(int kind, int n, int nz, bool fp32, int max_pn, int variant, int patch_elems) {
  return patch_dynamic_lds_bytes(kind, n, nz, fp32, max_pn, variant, patch_elems) + kStaticLds;
}

size_t patch_dynamic_lds_bytes(int kind, int n, int nz, bool fp32, int max_pn, int variant, int patch_elems) {
  if (variant == 1)
    return kind == 1 ? shell_line_lds_bytes(n, fp32, max_pn, patch_elems)
                     : solid_line_lds_bytes(n, nz, fp32, max_pn, patch_elems);
  if (variant == 2) return shell_mfma_lds_bytes(n, max_pn);
  if (variant == 3) return solid_stage_lds_bytes(n, nz, fp32, max_pn, patch_elems);
  const int nf = kind == 1 ? 5 : 3, fd = kind == 1 ? 2 : 3;
  const int np = kind == 1 ? n * n : n * n * nz;
  const int P = kind == 1 ? 16 : solid_patch_elems(n, nz, fp32);
  const size_t ts = fp32 ? 4 : 8;
  return region2_bytes(max_pn, nf, ts * (size_t)P * fd * np) + ts * ((size_t)P * nf * np + n * n + nz * nz + np) + 16;
}

// ---------------------------------------------------------------------------
// flat shell, 5 DOF/node (u, phix, v, phiy, w)
// coefficient row: a11 a12 a16 a22 a26 a66 a44 a45 a55 b11 b12 b16 b22 b26 b66
//                  d11 d12 d16 d22 d26 d66 m11 I11
// ---------------------------------------------------------------------------
// FP32: at most 96 VGPRs, so that two 16-element patches share a CU (the gather and
// assembly latency of one hides behind the other's arithmetic)
template <int N, typename T, bool AFF, bool AK, bool GR>
__global__ __launch_bounds__(ShellP<N>::TH, sizeof(T) == 4 ? 5 : 1) void shell_patch_kernel(const double* __restrict__ gtab,
                                                                   PatchArgs a, UpdateArgs ua) {
  constexpr int NP = N * N, NF = 5, TH = ShellP<N>::TH;
  extern __shared__ double smem[];
  double* Up = smem;
  T* fl = (T*)smem;
  T* su = (T*)((char*)smem + region2_bytes(a.max_pn, NF, sizeof(T) * ShellP<N>::P * 2 * NP));
  T* sD = su + ShellP<N>::P * NF * NP;
  T* sW = sD + N * N;
  const int tid = threadIdx.x, b = xcd_block(blockIdx.x, gridDim.x);
  const int pn0 = a.pn_base[b], pn = a.pn_base[b + 1] - pn0;
  const int e0 = a.el0[b], ne = a.el0[b + 1] - e0;
  gather_patch<NF, TH, double, gather_per_thread(ShellP<N>::P * NP, TH)>(a, ua, Up, b, pn0, pn);
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
    const int ln = a.lmap[a.lm_base[b] + le * NP + p];
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
      T g[5];
      point_geom<T, 5>(a, e * NP + p, g);
      xx = g[0]; xy = g[1]; yx = g[2]; yy = g[3];
      W = g[4];
    }
    const double* m = a.mat + (int64_t)a.elem_mat[e] * a.n_coef;
    const double sc = (a.sscale && !GR) ? a.sscale[e] : 1.0;
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
    if (!AK && !GR && a.excite && a.ex_ptr) apply_excitation(a, ua, e, p, NF, out);
    if (!GR) {
#pragma unroll
      for (int f = 0; f < NF; ++f) sue[f * NP + p] = out[f];
    }
  }
  if constexpr (GR) {
    // per-point contributions, then one thread per element sums its points in order
    // (the flux loop ended with a barrier, so the whole dynamic LDS is free here)
    double* red = smem;
    double as = 0.0, am = 0.0;
    if (valid) {
      const double* m = a.mat + (int64_t)a.elem_mat[e] * a.n_coef;
      const double Wp = AFF ? gtab[N * N + i] * gtab[N * N + j] * a.geom[e * 5 + 4] : a.geom[(e * NP + p) * 5 + 4];
      const double mf[5] = {m[21], m[22], m[21], m[22], m[21]};
      const int64_t nd = a.pn_id[pn0 + a.lmap[a.lm_base[b] + le * NP + p]];
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        const int64_t d = nd * NF + f;
        const double lam = ua.lam[d], un = ua.Unew[d], u = ua.U[d], uo = ua.Uold[d];
        as += lam * (double)out[f];
        am += lam * Wp * mf[f] * ((un - 2.0 * u + uo) + ua.gdv[f] * (un - uo));
      }
    }
    red[2 * tid] = as;
    red[2 * tid + 1] = am;
    __syncthreads();
    if (valid && p == 0) {
      double s1 = 0.0, s2 = 0.0;
      for (int q = 0; q < NP; ++q) {
        s1 += red[2 * (tid + q)];
        s2 += red[2 * (tid + q) + 1];
      }
      ua.gs[e] += ua.gscale_s * s1;
      ua.gr[e] += ua.gscale_m * s2;
    }
    return;
  }
  __syncthreads();
  patch_assemble<NF, NP, TH, AK, T>(a, ua, su, b, pn0, pn);
}

// ---------------------------------------------------------------------------
// hexahedral solid, 3 DOF/node (u, v, w), one wavefront per element
// coefficient row: D11 D12 D13 D14 D22 D23 D24 D33 D34 D44 D55 D56 D66 rho
// ---------------------------------------------------------------------------
template <int N, int NZ, typename T, bool AFF, bool AK, bool GR>
__global__ __launch_bounds__((SolidP<N, NZ>::TH)) void solid_patch_kernel(const double* __restrict__ gtab,
                                                                         PatchArgs a, UpdateArgs ua) {
  constexpr int NP = N * N * NZ, NF = 3, TH = SolidP<N, NZ>::TH, P = SolidP<N, NZ>::P,
                PPL = SolidP<N, NZ>::PPL;
  extern __shared__ double smem[];
  double* Up = smem;
  T* fl = (T*)smem;
  T* su = (T*)((char*)smem + region2_bytes(a.max_pn, NF, sizeof(T) * P * 3 * NP));
  T* sD = su + P * NF * NP;
  T* sDz = sD + N * N;
  T* sW = sDz + NZ * NZ;
  const int tid = threadIdx.x, b = xcd_block(blockIdx.x, gridDim.x);
  const int pn0 = a.pn_base[b], pn = a.pn_base[b + 1] - pn0;
  const int e0 = a.el0[b], ne = a.el0[b + 1] - e0;
  gather_patch<NF, TH, double, gather_per_thread(P * NP, TH)>(a, ua, Up, b, pn0, pn);
  for (int t = tid; t < N * N; t += TH) sD[t] = (T)gtab[t];
  for (int t = tid; t < NZ * NZ; t += TH) sDz[t] = (T)gtab[N * N + t];
  for (int t = tid; t < NP; t += TH) {
    const double* w = gtab + N * N + NZ * NZ;
    const double* wz = w + N;
    sW[t] = (T)(w[t % N] * w[(t / N) % N] * wz[t / (N * N)]);
  }
  __syncthreads();
  const int le = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int lane = tid & 63;
  if (le < ne) {
    const int64_t e = e0 + le;
    T* sue = su + le * NF * NP;
    T* fle = fl + le * 3 * NP;
#pragma unroll
    for (int pp = 0; pp < PPL; ++pp) {
      const int p = lane + 64 * pp;
      if (p < NP) {
        const int ln = a.lmap[a.lm_base[b] + le * NP + p];
#pragma unroll
        for (int f = 0; f < NF; ++f) sue[f * NP + p] = (T)Up[ln * NF + f];
      }
    }
  }
  __syncthreads();  // the fluxes below overwrite the gathered patch nodes
  if (le < ne) {
    const int64_t e = e0 + le;
    T* sue = su + le * NF * NP;
    T* fle = fl + le * 3 * NP;
    // element-uniform material (and geometry when affine)
    const double* m = a.mat + (int64_t)a.elem_mat[e] * a.n_coef;
    const T scT = (T)((a.sscale && !GR) ? a.sscale[e] : 1.0);
    const T D11 = (T)m[0], D12 = (T)m[1], D13 = (T)m[2], D14 = (T)m[3], D22 = (T)m[4], D23 = (T)m[5],
            D24 = (T)m[6], D33 = (T)m[7], D34 = (T)m[8], D44 = (T)m[9], D55 = (T)m[10], D56 = (T)m[11],
            D66 = (T)m[12];
    T iJe[3][3], detJe = 0;
    if (AFF) {
      const double* g = a.geom + e * 10;
#pragma unroll
      for (int q = 0; q < 9; ++q) iJe[q / 3][q % 3] = (T)g[q];
      detJe = (T)g[9];
    }
    wave_sync();
    T sg[PPL][6];   // Sxx Syy Szz Sxy Syz Sxz
    T Wq[PPL];
    T iJq[AFF ? 1 : PPL][3][3];
#pragma unroll
    for (int pp = 0; pp < PPL; ++pp) {
      const int p = lane + 64 * pp;
      const int i = p % N, j = (p / N) % N, k = p / (N * N);
      if (p < NP) {
        T dr[3][3];  // dr[f][d] = d u_f / d xi_d
#pragma unroll
        for (int f = 0; f < 3; ++f) {
          T s1 = 0, s2 = 0, s3 = 0;
#pragma unroll
          for (int mm = 0; mm < N; ++mm) {
            s1 += sD[i * N + mm] * sue[f * NP + (k * N + j) * N + mm];
            s2 += sD[j * N + mm] * sue[f * NP + (k * N + mm) * N + i];
          }
#pragma unroll
          for (int mm = 0; mm < NZ; ++mm) s3 += sDz[k * NZ + mm] * sue[f * NP + (mm * N + j) * N + i];
          dr[f][0] = s1;
          dr[f][1] = s2;
          dr[f][2] = s3;
        }
#define IJ(d, c) (AFF ? iJe[d][c] : iJq[AFF ? 0 : pp][d][c])
        if (AFF) {
          Wq[pp] = sW[p] * detJe;
        } else {
          T g[10];
          point_geom<T, 10>(a, e * NP + p, g);
#pragma unroll
          for (int q = 0; q < 9; ++q) iJq[AFF ? 0 : pp][q / 3][q % 3] = g[q];
          Wq[pp] = g[9];
        }
        T gr[3][3];  // gr[f][c] = d u_f / d x_c (strains_spec_p.m:8-27)
#pragma unroll
        for (int f = 0; f < 3; ++f)
#pragma unroll
          for (int c = 0; c < 3; ++c) gr[f][c] = dr[f][0] * IJ(0, c) + dr[f][1] * IJ(1, c) + dr[f][2] * IJ(2, c);
        const T Exx = gr[0][0], Eyy = gr[1][1], Ezz = gr[2][2];
        const T Exy = gr[0][1] + gr[1][0], Eyz = gr[1][2] + gr[2][1], Exz = gr[0][2] + gr[2][0];
        // stresses_spec_p.m:6-11
        sg[pp][0] = scT * (Exx * D11 + Eyy * D12 + Ezz * D13 + Exy * D14);
        sg[pp][1] = scT * (Exx * D12 + Eyy * D22 + Ezz * D23 + Exy * D24);
        sg[pp][2] = scT * (Exx * D13 + Eyy * D23 + Ezz * D33 + Exy * D34);
        sg[pp][3] = scT * (Exx * D14 + Eyy * D24 + Ezz * D34 + Exy * D44);
        sg[pp][4] = scT * (Eyz * D55 + Exz * D56);
        sg[pp][5] = scT * (Eyz * D56 + Exz * D66);
      }
    }
    T out[PPL][3];
#pragma unroll
    for (int pp = 0; pp < PPL; ++pp) out[pp][0] = out[pp][1] = out[pp][2] = 0;
    // transposed derivatives, one reference direction at a time:
    // f_u += D_d^T (W sum_c sig[u][c] iJ[d][c])  (main_spec3Dcuda_global_..._mat.m:944-954)
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      wave_sync();
#pragma unroll
      for (int pp = 0; pp < PPL; ++pp) {
        const int p = lane + 64 * pp;
        if (p < NP) {
          const T W = Wq[pp];
          fle[0 * NP + p] = W * (sg[pp][0] * IJ(d, 0) + sg[pp][3] * IJ(d, 1) + sg[pp][5] * IJ(d, 2));
          fle[1 * NP + p] = W * (sg[pp][3] * IJ(d, 0) + sg[pp][1] * IJ(d, 1) + sg[pp][4] * IJ(d, 2));
          fle[2 * NP + p] = W * (sg[pp][5] * IJ(d, 0) + sg[pp][4] * IJ(d, 1) + sg[pp][2] * IJ(d, 2));
        }
      }
      wave_sync();
#pragma unroll
      for (int pp = 0; pp < PPL; ++pp) {
        const int p = lane + 64 * pp;
        const int i = p % N, j = (p / N) % N, k = p / (N * N);
        if (p < NP) {
#pragma unroll
          for (int f = 0; f < 3; ++f) {
            T s = 0;
            if (d == 0) {
#pragma unroll
              for (int mm = 0; mm < N; ++mm) s += sD[mm * N + i] * fle[f * NP + (k * N + j) * N + mm];
            } else if (d == 1) {
#pragma unroll
              for (int mm = 0; mm < N; ++mm) s += sD[mm * N + j] * fle[f * NP + (k * N + mm) * N + i];
            } else {
#pragma unroll
              for (int mm = 0; mm < NZ; ++mm) s += sDz[mm * NZ + k] * fle[f * NP + (mm * N + j) * N + i];
            }
            out[pp][f] += s;
          }
        }
      }
    }
#undef IJ
    wave_sync();
#pragma unroll
    for (int pp = 0; pp < PPL; ++pp) {
      const int p = lane + 64 * pp;
      if (p < NP) {
        if (!AK && !GR && a.excite && a.ex_ptr) apply_excitation(a, ua, e, p, NF, out[pp]);
        if (!GR) {
#pragma unroll
          for (int f = 0; f < NF; ++f) sue[f * NP + p] = out[pp][f];
        }
      }
    }
    if constexpr (GR) {
      double as = 0.0, am = 0.0;
      const double rho = m[13];
      const double* w = gtab + N * N + NZ * NZ;
#pragma unroll
      for (int pp = 0; pp < PPL; ++pp) {
        const int p = lane + 64 * pp;
        if (p < NP) {
          const int64_t nd = a.pn_id[pn0 + a.lmap[a.lm_base[b] + le * NP + p]];
          const double Wp = AFF ? w[p % N] * w[(p / N) % N] * w[N + p / (N * N)] * a.geom[e * 10 + 9]
                                : a.geom[(e * NP + p) * 10 + 9];
#pragma unroll
          for (int f = 0; f < NF; ++f) {
            const int64_t d = nd * NF + f;
            const double lam = ua.lam[d], un = ua.Unew[d], u = ua.U[d], uo = ua.Uold[d];
            as += lam * (double)out[pp][f];
            am += lam * Wp * rho * ((un - 2.0 * u + uo) + ua.gdv[f] * (un - uo));
          }
        }
      }
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) {
        as += __shfl_down(as, off, 64);
        am += __shfl_down(am, off, 64);
      }
      if (lane == 0) {
        ua.gs[e] += ua.gscale_s * as;
        ua.gr[e] += ua.gscale_m * am;
      }
    }
  }
  if constexpr (GR) return;
  __syncthreads();
  patch_assemble<NF, NP, TH, AK, T>(a, ua, su, b, pn0, pn);
}

// ---------------------------------------------------------------------------
// patch-boundary nodes: partial sums + update + their sensor terms
// ---------------------------------------------------------------------------
// T: the arithmetic type of the patch kernel that wrote the partials.
template <int NF, bool AK, typename T>
__global__ __launch_bounds__(256) void shared_update_kernel(SharedArgs s, UpdateArgs ua) {
  const int64_t k = (int64_t)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (k >= s.n_shared) return;
  const uint64_t nn = AK ? 0 : ua.ctr[0] + ua.step_off;
  const int fi = AK ? -1 : frame_of(ua, nn);
  bool bad = false;
  const int64_t nd = s.sh_id[k];
  double F[NF];
#pragma unroll
  for (int f = 0; f < NF; ++f) F[f] = 0.0;
  const T* part = (const T*)ua.partial;
  // the first KS partials load independently (indices clamped to the node's last slot, a
  // node has at least one), so their latencies overlap; summed in patch order
  constexpr int KS = 4;
  const int o0 = s.sh_ptr[k], o1 = s.sh_ptr[k + 1];
  T v[KS][NF];
#pragma unroll
  for (int i = 0; i < KS; ++i) {
    const int64_t sl = s.sh_slot[o0 + i < o1 ? o0 + i : o1 - 1];
#pragma unroll
    for (int f = 0; f < NF; ++f) v[i][f] = part[sl * NF + f];
  }
#pragma unroll
  for (int i = 0; i < KS; ++i)
    if (o0 + i < o1) {
#pragma unroll
      for (int f = 0; f < NF; ++f) F[f] += (double)v[i][f];
    }
  for (int o = o0 + KS; o < o1; ++o) {
    const T* src = part + (int64_t)s.sh_slot[o] * NF;
#pragma unroll
    for (int f = 0; f < NF; ++f) F[f] += (double)src[f];
  }
  NodeState<NF, T> st;
  if (!AK) {
    st.load_u(ua, nd);
    st.load_rest(ua, nd);
  }
  finish_node<NF, AK, T>(ua, nd, F, st, fi, bad, ArgConsts{ua});
  if (bad) atomicMin((unsigned long long*)&ua.ctr[1], (unsigned long long)nn);
  if (!AK && s.ts_ptr) {
    for (int q = s.ts_ptr[k]; q < s.ts_ptr[k + 1]; ++q) {
      const int f = s.ts_f[q];
      const int64_t d = nd * NF + f;
      ua.ring[ua.step_off * ua.n_terms + s.ts_t[q]] = sample(ua.out_q, ua.U[d], ua.Uold[d], ua.Unew[d], ua.a0, ua.a1);
    }
  }
}

// ---------------------------------------------------------------------------
// end of a chunk of steps: output functionals from the ring, step counter
// ---------------------------------------------------------------------------
// One wavefront per (step, output): lane-strided partial sums over the output's terms,
// then a fixed shuffle tree (run-to-run reproducible).
__global__ __launch_bounds__(64) void output_kernel(AdvanceArgs a) {
  const int st = blockIdx.x / a.n_out, jo = blockIdx.x - st * a.n_out;
  const uint64_t nn0 = a.ctr[0];
  const int64_t t0 = a.out_ptr[jo], t1 = a.out_ptr[jo + 1];
  double acc = 0.0;
  for (int64_t t = t0 + threadIdx.x; t < t1; t += 64) acc += a.out_w[t] * a.ring[st * a.n_terms + t];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  if (threadIdx.x == 0) a.traces[jo * a.ld + ((int64_t)nn0 + st - a.nn_base)] = acc;
}

__global__ void counter_kernel(uint64_t* ctr, int count) { ctr[0] += (uint64_t)count; }

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
template <int N, typename T, bool AFF, bool AK, bool GR>
static hipError_t shell_go(int64_t n_patch, const double* gtab, const PatchArgs& pa, const UpdateArgs& ua,
                           hipStream_t s) {
  const size_t lds = patch_dynamic_lds_bytes(1, N, 1, sizeof(T) == 4, pa.max_pn, 0, 16);
  static LdsGrant granted;
  hipError_t e = lds_opt_in(shell_patch_kernel<N, T, AFF, AK, GR>, lds, granted);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((shell_patch_kernel<N, T, AFF, AK, GR>), dim3((unsigned)n_patch), dim3(ShellP<N>::TH), lds, s,
                     gtab, pa, ua);
  return hipGetLastError();
}

template <int N, int NZ, typename T, bool AFF, bool AK, bool GR>
static hipError_t solid_go(int64_t n_patch, const double* gtab, const PatchArgs& pa, const UpdateArgs& ua,
                           hipStream_t s) {
  const size_t lds = patch_dynamic_lds_bytes(2, N, NZ, sizeof(T) == 4, pa.max_pn, 0, 4);
  static LdsGrant granted;
  hipError_t e = lds_opt_in(solid_patch_kernel<N, NZ, T, AFF, AK, GR>, lds, granted);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((solid_patch_kernel<N, NZ, T, AFF, AK, GR>), dim3((unsigned)n_patch), dim3(SolidP<N, NZ>::TH),
                     lds, s, gtab, pa, ua);
  return hipGetLastError();
}

template <int N, int NZ, bool SHELL, typename T, bool A>
static hipError_t dispatch_mode(int mode, int64_t np, const double* g, const PatchArgs& pa, const UpdateArgs& ua,
                                hipStream_t s) {
  if constexpr (SHELL) {
    if (mode == 0) return shell_go<N, T, A, false, false>(np, g, pa, ua, s);
    if (mode == 1) return shell_go<N, T, A, true, false>(np, g, pa, ua, s);
    return shell_go<N, T, A, true, true>(np, g, pa, ua, s);
  } else {
    if (mode == 0) return solid_go<N, NZ, T, A, false, false>(np, g, pa, ua, s);
    if (mode == 1) return solid_go<N, NZ, T, A, true, false>(np, g, pa, ua, s);
    return solid_go<N, NZ, T, A, true, true>(np, g, pa, ua, s);
  }
}

template <int N, int NZ, bool SHELL>
static hipError_t dispatch(bool fp32, bool aff, int mode, int64_t np, const double* g, const PatchArgs& pa,
                           const UpdateArgs& ua, hipStream_t s) {
  if (fp32) return aff ? dispatch_mode<N, NZ, SHELL, float, true>(mode, np, g, pa, ua, s)
                       : dispatch_mode<N, NZ, SHELL, float, false>(mode, np, g, pa, ua, s);
  return aff ? dispatch_mode<N, NZ, SHELL, double, true>(mode, np, g, pa, ua, s)
             : dispatch_mode<N, NZ, SHELL, double, false>(mode, np, g, pa, ua, s);
}

hipError_t launch_patch(int kind, int n, int nz, bool fp32, bool affine, int mode, int variant, int patch_elems,
                        int64_t n_patch, const double* gtab, const PatchArgs& pa, const UpdateArgs& ua,
                        hipStream_t s) {
  if (!variant_accepts(kind, variant, patch_elems)) return hipErrorInvalidValue;
  if (variant == 1)
    return kind == 1 ? launch_shell_line(n, fp32, affine, mode, patch_elems, n_patch, gtab, pa, ua, s)
                     : launch_solid_line(n, nz, fp32, affine, mode, patch_elems, n_patch, gtab, pa, ua, s);
  if (variant == 2) return fp32 ? launch_shell_mfma(n, affine, mode, n_patch, gtab, pa, ua, s) : hipErrorInvalidValue;
  if (variant == 3) return launch_solid_stage(n, nz, fp32, affine, mode, patch_elems, n_patch, gtab, pa, ua, s);
  if (kind == 1) {
    switch (n) {
      case 3: return dispatch<3, 1, true>(fp32, affine, mode, n_patch, gtab, pa, ua, s);
      case 4: return dispatch<4, 1, true>(fp32, affine, mode, n_patch, gtab, pa, ua, s);
      case 5: return dispatch<5, 1, true>(fp32, affine, mode, n_patch, gtab, pa, ua, s);
      case 6: return dispatch<6, 1, true>(fp32, affine, mode, n_patch, gtab, pa, ua, s);
      case 7: return dispatch<7, 1, true>(fp32, affine, mode, n_patch, gtab, pa, ua, s);
      case 8: return dispatch<8, 1, true>(fp32, affine, mode, n_patch, gtab, pa, ua, s);
      default: return hipErrorInvalidValue;
    }
  }
#define SOLID_CASE(A, B) \
  if (n == A && nz == B) return dispatch<A, B, false>(fp32, affine, mode, n_patch, gtab, pa, ua, s);
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

hipError_t launch_shared(int nf, bool applyK, bool fp32, const SharedArgs& sa, const UpdateArgs& ua,
                         hipStream_t s) {
  if (sa.n_shared <= 0) return hipSuccess;
  const unsigned blocks = (unsigned)((sa.n_shared + 255) / 256);
#define SH(NFV, AK, T) hipLaunchKernelGGL((shared_update_kernel<NFV, AK, T>), dim3(blocks), dim3(256), 0, s, sa, ua)
#define SH2(NFV, AK) { if (fp32) SH(NFV, AK, float); else SH(NFV, AK, double); }
  if (nf == 5) { if (applyK) SH2(5, true) else SH2(5, false) }
  else if (nf == 3) { if (applyK) SH2(3, true) else SH2(3, false) }
  else return hipErrorInvalidValue;
#undef SH2
#undef SH
  return hipGetLastError();
}

hipError_t launch_advance(const AdvanceArgs& a, hipStream_t s) {
  if (a.count > 0 && a.n_out > 0) {
    hipLaunchKernelGGL(output_kernel, dim3((unsigned)(a.count * a.n_out)), dim3(64), 0, s, a);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(counter_kernel, dim3(1), dim3(1), 0, s, a.ctr, a.count);
  return hipGetLastError();
}

}  // namespace semk
