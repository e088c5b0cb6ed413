// Hexahedral solid patch kernel with line-wise GLL contractions (gfx950).
//
// One wavefront per element, P elements (one patch) per workgroup, as the point-wise
// kernel; the difference is how the sum-factorised contractions are mapped.  A lane owns
// a whole GLL line of one field (xi, eta or zeta direction): it loads the line's N values
// from LDS once and forms all N derivatives along it with the derivative matrix held in
// scalar registers (wavefront-uniform), so each LDS value is read once per direction
// instead of once per multiply-add.  The point-wise stage (strains, constitutive product,
// fluxes) then handles one point at a time, which keeps register use low at every order.
//
// LDS per element: derivatives/fluxes [3 dir][3 field][NP], which the patch-node gather
// region aliases; with N == NZ the element values and forces [3][NP] live inside the
// element's own derivative region (values in the zeta blocks, forces in the xi blocks),
// otherwise in a region of their own.
#include "sem_device.hpp"

namespace semk {

namespace {

// Where point p = i + N j + N^2 k sits in each direction's block of derivatives: LM stores
// the lines of a direction one after the other (xi line (j, k) at (k N + j) N, eta line
// (i, k) at (k N + i) N, zeta line (i, j) at (j N + i) N, entry along the line last);
// otherwise every block is in point order.
template <int N, bool LM>
__device__ __forceinline__ void line_index(int p, int* ix) {
  if constexpr (LM) {
    const int i = p % N, jk = p / N, j = jk % N, k = jk / N;
    ix[0] = p;
    ix[1] = (k * N + i) * N + j;
    ix[2] = (j * N + i) * N + k;
  } else {
    ix[0] = ix[1] = ix[2] = p;
  }
}

template <bool LM>
__device__ __forceinline__ void point_index(const int* sPt, int p, int* ix) {
  ix[0] = p;
  if constexpr (LM) {
    const int w = sPt[p];
    ix[1] = w & 0xffff;
    ix[2] = w >> 16;
  } else {
    ix[1] = ix[2] = p;
  }
}

}  // namespace

size_t solid_line_lds_bytes(int n, int nz, bool fp32, int max_pn, int P) {
  const size_t ts = fp32 ? 4 : 8, np = (size_t)n * n * nz;
  const size_t tables = n == nz ? 4 * (9 * (size_t)n * n + np) : 0, values = n == nz ? 0 : (size_t)P * 3 * np;
  return region2_bytes(max_pn, 3, ts * (size_t)P * 9 * np) + ts * (values + np) + tables + 16;
}

template <int N, int NZ, typename T, bool AFF, bool AK, bool GR, int P>
__global__ __launch_bounds__(64 * P, sizeof(T) == 4 ? 8 : 1) void solid_line_kernel(const double* __restrict__ gtab, PatchArgs a,
                                                            UpdateArgs ua) {
  constexpr int NP = N * N * NZ, NF = 3, TH = 64 * P, NLX = N * NZ, NLZ = N * N;
  constexpr bool LM = N == NZ;
  extern __shared__ double smem[];
  double* Up = smem;
  T* dv = (T*)smem;  // [P][3 dir][3 field][NP]
  T* su = (T*)((char*)smem + region2_bytes(a.max_pn, NF, sizeof(T) * (size_t)P * 9 * NP));  // [P][3][NP], !LM
  T* sW = LM ? su : su + P * NF * NP;  // [NP] quadrature weight products w_i w_j w_k
  // LM tables: per line task, (offset of its first value in an element's [3][NP] values) |
  // (stride along the line) << 16; per point, (eta index) | (zeta index) << 16
  int* sLine = (int*)(sW + NP);
  int* sPt = sLine + 9 * NLZ;
  const DerivTab<T, N, NZ> tb{gtab};
  const int tid = threadIdx.x, b = xcd_block(blockIdx.x, gridDim.x);
  const int pn0 = a.pn_base[b], pn = a.pn_base[b + 1] - pn0;
  const int e0 = a.el0[b], ne = a.el0[b + 1] - e0;
  gather_patch<NF, TH, double, gather_per_thread(P * NP, TH)>(a, ua, Up, b, pn0, pn);
  OccHeld<2> occ;  // assembly indices of this thread's first two patch nodes
  if constexpr (!GR) occ_prefetch<TH, 2>(a, b, pn0, pn, occ);
  if constexpr (AFF) {
    const double* w = gtab + N * N + NZ * NZ;
    for (int t = tid; t < NP; t += TH) sW[t] = (T)(w[t % N] * w[(t / N) % N] * w[N + t / (N * N)]);
  }
  if constexpr (LM) {
    for (int t = tid; t < 9 * NLZ; t += TH) {
      const int df = t / NLZ, r = t - df * NLZ, d = df / 3, f = df - 3 * d;
      const int q = r / N, s = r - q * N;
      // xi: line (j = s, k = q); eta: line (i = s, k = q); zeta: line (i = s, j = q)
      const int base = d == 0 ? r * N : (d == 1 ? q * N * N + s : r);
      const int stride = d == 0 ? 1 : (d == 1 ? N : N * N);
      sLine[t] = (f * NP + base) | (stride << 16);
    }
    for (int t = tid; t < NP; t += TH) {
      int ix[3];
      line_index<N, LM>(t, ix);
      sPt[t] = ix[1] | (ix[2] << 16);
    }
  }
  __syncthreads();
  const int le = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int lane = tid & 63;
  const int64_t e = e0 + le;
  T* dve = dv + le * 9 * NP;
  // element values [3][NP]: LM in the zeta blocks of the element's derivative region, so
  // they pass through registers (the patch nodes they come from share that region)
  T* sue = LM ? dve + 6 * NP : su + le * NF * NP;
  // element forces [3][NP]: LM in the xi blocks (written in place by the last pass)
  T* fe = LM ? dve : sue;
  if constexpr (LM) {
    constexpr int KP = (NP + 63) / 64;
    T v[KP][NF];
    if (le < ne) {
#pragma unroll
      for (int k = 0; k < KP; ++k) {
        const int p = lane + 64 * k;
        if (p < NP) {
          const int ln = a.lmap[a.lm_base[b] + le * NP + p];
#pragma unroll
          for (int f = 0; f < NF; ++f) v[k][f] = (T)Up[ln * NF + f];
        }
      }
    }
    __syncthreads();  // all patch-node reads done before the element regions are written
    if (le < ne) {
#pragma unroll
      for (int k = 0; k < KP; ++k) {
        const int p = lane + 64 * k;
        if (p < NP) {
#pragma unroll
          for (int f = 0; f < NF; ++f) sue[f * NP + p] = v[k][f];
        }
      }
    }
    wave_sync();
  } else {
    if (le < ne) {
      for (int p = lane; p < NP; p += 64) {
        const int ln = a.lmap[a.lm_base[b] + le * NP + p];
#pragma unroll
        for (int f = 0; f < NF; ++f) sue[f * NP + p] = (T)Up[ln * NF + f];
      }
    }
    __syncthreads();  // the derivatives below overwrite the gathered patch nodes
  }
  if (le < ne) {
    // --- forward derivatives, one line per lane-task.  With N == NZ the 9 (direction, field)
    // blocks of N*N lines are one task list (4 passes of 64 lanes instead of 6 at N = 5), and
    // task t stores its N derivatives at dve[t*N ..] (line-major: LM); otherwise one loop per
    // direction stores them at dve[(d*3 + f)*NP + p].
    if constexpr (LM) {
      // Tasks t < 6 N^2 write below the element values; the passes that hold later tasks
      // (which overwrite the values) load all their lines before any of them stores.
      constexpr int NT = 9 * NLZ, NIT = (NT + 63) / 64, IS = 6 * NLZ / 64, NS = NIT - IS;
      auto contract = [&](int t, const T* u) {
        T* dst = dve + t * N;
#pragma unroll
        for (int i = 0; i < N; ++i) {
          T s2 = 0;
#pragma unroll
          for (int m = 0; m < N; ++m) s2 += tb.D(i, m) * u[m];
          dst[i] = s2;
        }
      };
      for (int it = 0; it < IS; ++it) {
        const int t = lane + 64 * it;
        const int w = sLine[t], stride = w >> 16;
        const T* src = sue + (w & 0xffff);
        T u[N];
#pragma unroll
        for (int m = 0; m < N; ++m) u[m] = src[m * stride];
        contract(t, u);
      }
      T us[NS][N];
#pragma unroll
      for (int k = 0; k < NS; ++k) {
        const int t = lane + 64 * (IS + k);
        if (t < NT) {
          const int w = sLine[t], stride = w >> 16;
          const T* src = sue + (w & 0xffff);
#pragma unroll
          for (int m = 0; m < N; ++m) us[k][m] = src[m * stride];
        }
      }
      wave_sync();
#pragma unroll
      for (int k = 0; k < NS; ++k) {
        const int t = lane + 64 * (IS + k);
        if (t < NT) contract(t, us[k]);
      }
    } else {
      for (int t = lane; t < NF * NLX; t += 64) {  // xi lines (j, k)
        const int f = t / NLX, r = t - f * NLX;
        const T* src = sue + f * NP + r * N;
        T u[N];
#pragma unroll
        for (int m = 0; m < N; ++m) u[m] = src[m];
        T* dst = dve + f * NP + r * N;
#pragma unroll
        for (int i = 0; i < N; ++i) {
          T s = 0;
#pragma unroll
          for (int m = 0; m < N; ++m) s += tb.D(i, m) * u[m];
          dst[i] = s;
        }
      }
      for (int t = lane; t < NF * NLX; t += 64) {  // eta lines (i, k)
        const int f = t / NLX, r = t - f * NLX, k = r / N, i = r - k * N;
        const T* src = sue + f * NP + k * N * N + i;
        T u[N];
#pragma unroll
        for (int m = 0; m < N; ++m) u[m] = src[m * N];
        T* dst = dve + (3 + f) * NP + k * N * N + i;
#pragma unroll
        for (int j = 0; j < N; ++j) {
          T s = 0;
#pragma unroll
          for (int m = 0; m < N; ++m) s += tb.D(j, m) * u[m];
          dst[j * N] = s;
        }
      }
      for (int t = lane; t < NF * NLZ; t += 64) {  // zeta lines (i, j)
        const int f = t / NLZ, r = t - f * NLZ;
        const T* src = sue + f * NP + r;
        T u[NZ];
#pragma unroll
        for (int m = 0; m < NZ; ++m) u[m] = src[m * N * N];
        T* dst = dve + (6 + f) * NP + r;
#pragma unroll
        for (int k = 0; k < NZ; ++k) {
          T s = 0;
#pragma unroll
          for (int m = 0; m < NZ; ++m) s += tb.Dz(k, m) * u[m];
          dst[k * N * N] = s;
        }
      }
    }
    wave_sync();
    // --- point-wise: strains, stresses (stresses_spec_p.m:6-11), fluxes in place
    const double* m = a.mat + (int64_t)a.elem_mat[e] * a.n_coef;
    const double sc = (a.sscale && !GR) ? a.sscale[e] : 1.0;
    const T D11 = (T)(m[0] * sc), D12 = (T)(m[1] * sc), D13 = (T)(m[2] * sc), D14 = (T)(m[3] * sc),
            D22 = (T)(m[4] * sc), D23 = (T)(m[5] * sc), D24 = (T)(m[6] * sc), D33 = (T)(m[7] * sc),
            D34 = (T)(m[8] * sc), D44 = (T)(m[9] * sc), D55 = (T)(m[10] * sc), D56 = (T)(m[11] * sc),
            D66 = (T)(m[12] * sc);
    T iJe[3][3], We = 0;
    if constexpr (AFF) {
      const double* g = a.geom + e * 10;
#pragma unroll
      for (int q = 0; q < 9; ++q) iJe[q / 3][q % 3] = (T)g[q];
      We = (T)g[9];
    }
    for (int p = lane; p < NP; p += 64) {
      T iJ[3][3], W;
      if constexpr (AFF) {
#pragma unroll
        for (int q = 0; q < 9; ++q) iJ[q / 3][q % 3] = iJe[q / 3][q % 3];
        W = sW[p] * We;
      } else {
        T g[10];
        point_geom<T, 10>(a, e * NP + p, g);
#pragma unroll
        for (int q = 0; q < 9; ++q) iJ[q / 3][q % 3] = g[q];
        W = g[9];
      }
      int ix[3];
      point_index<LM>(sPt, p, ix);
      T gr[3][3];  // gr[f][c] = d u_f / d x_c (strains_spec_p.m:8-27)
#pragma unroll
      for (int f = 0; f < 3; ++f) {
        const T d0 = dve[f * NP + ix[0]], d1 = dve[(3 + f) * NP + ix[1]], d2 = dve[(6 + f) * NP + ix[2]];
#pragma unroll
        for (int c = 0; c < 3; ++c) gr[f][c] = d0 * iJ[0][c] + d1 * iJ[1][c] + d2 * iJ[2][c];
      }
      const T Exx = gr[0][0], Eyy = gr[1][1], Ezz = gr[2][2];
      const T Exy = gr[0][1] + gr[1][0], Eyz = gr[1][2] + gr[2][1], Exz = gr[0][2] + gr[2][0];
      const T Sxx = Exx * D11 + Eyy * D12 + Ezz * D13 + Exy * D14;
      const T Syy = Exx * D12 + Eyy * D22 + Ezz * D23 + Exy * D24;
      const T Szz = Exx * D13 + Eyy * D23 + Ezz * D33 + Exy * D34;
      const T Sxy = Exx * D14 + Eyy * D24 + Ezz * D34 + Exy * D44;
      const T Syz = Eyz * D55 + Exz * D56;
      const T Sxz = Eyz * D56 + Exz * D66;
      // f_u += D_d^T (W sum_c sig[u][c] iJ[d][c])  (main_spec3Dcuda_global_..._mat.m:944-954)
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        dve[(3 * d + 0) * NP + ix[d]] = W * (Sxx * iJ[d][0] + Sxy * iJ[d][1] + Sxz * iJ[d][2]);
        dve[(3 * d + 1) * NP + ix[d]] = W * (Sxy * iJ[d][0] + Syy * iJ[d][1] + Syz * iJ[d][2]);
        dve[(3 * d + 2) * NP + ix[d]] = W * (Sxz * iJ[d][0] + Syz * iJ[d][1] + Szz * iJ[d][2]);
      }
    }
    wave_sync();
    // --- transposed derivatives.  LM: every line in place (4 passes of 64 lanes), then the
    // element forces per point as (xi + eta) + zeta; otherwise xi writes the element forces
    // and eta and zeta add to them (the same sums in the same order).
    if constexpr (LM) {
      for (int t = lane; t < 9 * NLZ; t += 64) {
        T* v = dve + t * N;
        T u[N];
#pragma unroll
        for (int mm = 0; mm < N; ++mm) u[mm] = v[mm];
#pragma unroll
        for (int i = 0; i < N; ++i) {
          T s2 = 0;
#pragma unroll
          for (int mm = 0; mm < N; ++mm) s2 += tb.D(mm, i) * u[mm];
          v[i] = s2;
        }
      }
      wave_sync();
      for (int p = lane; p < NP; p += 64) {
        int ix[3];
        point_index<LM>(sPt, p, ix);
#pragma unroll
        for (int f = 0; f < NF; ++f)
          fe[f * NP + p] = (dve[f * NP + ix[0]] + dve[(3 + f) * NP + ix[1]]) + dve[(6 + f) * NP + ix[2]];
      }
      wave_sync();
    } else {
      for (int t = lane; t < NF * NLX; t += 64) {
        const int f = t / NLX, r = t - f * NLX;
        const T* src = dve + f * NP + r * N;
        T u[N];
#pragma unroll
        for (int mm = 0; mm < N; ++mm) u[mm] = src[mm];
        T* dst = sue + f * NP + r * N;
#pragma unroll
        for (int i = 0; i < N; ++i) {
          T s = 0;
#pragma unroll
          for (int mm = 0; mm < N; ++mm) s += tb.D(mm, i) * u[mm];
          dst[i] = s;
        }
      }
      wave_sync();
      for (int t = lane; t < NF * NLX; t += 64) {
        const int f = t / NLX, r = t - f * NLX, k = r / N, i = r - k * N;
        const T* src = dve + (3 + f) * NP + k * N * N + i;
        T u[N];
#pragma unroll
        for (int mm = 0; mm < N; ++mm) u[mm] = src[mm * N];
        T* dst = sue + f * NP + k * N * N + i;
#pragma unroll
        for (int j = 0; j < N; ++j) {
          T s = 0;
#pragma unroll
          for (int mm = 0; mm < N; ++mm) s += tb.D(mm, j) * u[mm];
          dst[j * N] += s;
        }
      }
      wave_sync();
      for (int t = lane; t < NF * NLZ; t += 64) {
        const int f = t / NLZ, r = t - f * NLZ;
        const T* src = dve + (6 + f) * NP + r;
        T u[NZ];
#pragma unroll
        for (int mm = 0; mm < NZ; ++mm) u[mm] = src[mm * N * N];
        T* dst = sue + f * NP + r;
#pragma unroll
        for (int k = 0; k < NZ; ++k) {
          T s = 0;
#pragma unroll
          for (int mm = 0; mm < NZ; ++mm) s += tb.Dz(mm, k) * u[mm];
          dst[k * N * N] += s;
        }
      }
      wave_sync();
    }
    // --- point sources (serial per element: entries may share a point)
    if constexpr (!AK && !GR) {
      if (a.excite && a.ex_ptr && lane == 0) {
        const int b0 = a.ex_ptr[e], b1 = a.ex_ptr[e + 1];
        if (b0 < b1) {
          const int64_t col = (int64_t)(ua.ctr[0] + ua.step_off) - 1;
          for (int q = b0; q < b1; ++q) {
            const int pf = a.ex_pf[q];
            fe[(pf % NF) * NP + pf / NF] -= (T)(a.ex_amp[q] * a.st[(int64_t)a.ex_sig[q] * a.ex_nft + col]);
          }
        }
      }
    }
    if constexpr (GR) {
      double as = 0.0, am = 0.0;
      const double rho = m[13];
      const double* w = gtab + N * N + NZ * NZ;
      for (int p = lane; p < NP; p += 64) {
        const int64_t nd = a.pn_id[pn0 + a.lmap[a.lm_base[b] + le * NP + p]];
        const double Wp = AFF ? w[p % N] * w[(p / N) % N] * w[N + p / (N * N)] * a.geom[e * 10 + 9]
                              : a.geom[(e * NP + p) * 10 + 9];
#pragma unroll
        for (int f = 0; f < NF; ++f) {
          const int64_t d = nd * NF + f;
          const double lam = ua.lam[d], un = ua.Unew[d], u = ua.U[d], uo = ua.Uold[d];
          as += lam * (double)fe[f * NP + p];
          am += lam * Wp * rho * ((un - 2.0 * u + uo) + ua.gdv[f] * (un - uo));
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
  if constexpr (LM) patch_assemble<NF, NP, TH, AK, T, 9 * NP, 2>(a, ua, dv, b, pn0, pn, &occ);
  else patch_assemble<NF, NP, TH, AK, T, NF * NP, 2>(a, ua, su, b, pn0, pn, &occ);
}

namespace {

template <int N, int NZ, typename T, bool AFF, bool AK, bool GR, int P>
hipError_t go(int64_t n_patch, const double* gtab, const PatchArgs& pa, const UpdateArgs& ua, hipStream_t s) {
  const size_t lds = solid_line_lds_bytes(N, NZ, sizeof(T) == 4, pa.max_pn, P);
  static LdsGrant granted;
  hipError_t e = lds_opt_in(solid_line_kernel<N, NZ, T, AFF, AK, GR, P>, lds, granted);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((solid_line_kernel<N, NZ, T, AFF, AK, GR, P>), dim3((unsigned)n_patch), dim3(64 * P), lds, s,
                     gtab, pa, ua);
  return hipGetLastError();
}

template <int N, int NZ, typename T, bool A, int P>
hipError_t by_mode(int mode, int64_t np, const double* g, const PatchArgs& pa, const UpdateArgs& ua, hipStream_t s) {
  if (mode == 0) return go<N, NZ, T, A, false, false, P>(np, g, pa, ua, s);
  if (mode == 1) return go<N, NZ, T, A, true, false, P>(np, g, pa, ua, s);
  return go<N, NZ, T, A, true, true, P>(np, g, pa, ua, s);
}

template <int N, int NZ>
hipError_t by_type(bool fp32, bool aff, int mode, int P, int64_t np, const double* g, const PatchArgs& pa,
                   const UpdateArgs& ua, hipStream_t s) {
  if (P == 8) {
    if (fp32) return aff ? by_mode<N, NZ, float, true, 8>(mode, np, g, pa, ua, s)
                         : by_mode<N, NZ, float, false, 8>(mode, np, g, pa, ua, s);
    return aff ? by_mode<N, NZ, double, true, 8>(mode, np, g, pa, ua, s)
               : by_mode<N, NZ, double, false, 8>(mode, np, g, pa, ua, s);
  }
  if (P == 4) {
    if (fp32) return aff ? by_mode<N, NZ, float, true, 4>(mode, np, g, pa, ua, s)
                         : by_mode<N, NZ, float, false, 4>(mode, np, g, pa, ua, s);
    return aff ? by_mode<N, NZ, double, true, 4>(mode, np, g, pa, ua, s)
               : by_mode<N, NZ, double, false, 4>(mode, np, g, pa, ua, s);
  }
  return hipErrorInvalidValue;
}

}  // namespace

hipError_t launch_solid_line(int n, int nz, bool fp32, bool affine, int mode, int P, int64_t n_patch,
                             const double* gtab, const PatchArgs& pa, const UpdateArgs& ua, hipStream_t s) {
#define LINE_CASE(A, B) \
  if (n == A && nz == B) return by_type<A, B>(fp32, affine, mode, P, n_patch, gtab, pa, ua, s);
  LINE_CASE(3, 3)
  LINE_CASE(4, 4)
  LINE_CASE(4, 3)
  LINE_CASE(5, 5)
  LINE_CASE(5, 3)
  LINE_CASE(6, 6)
  LINE_CASE(6, 3)
#undef LINE_CASE
  return hipErrorInvalidValue;
}

}  // namespace semk
