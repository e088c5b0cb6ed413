// Flat-shell patch kernel with line-wise GLL contractions (gfx950).
//
// One wavefront per element, P elements (one patch) per workgroup.  A lane owns one GLL
// line of one field in one direction, 2 x 5 x N lines per element (60 at order 5: one pass
// of 64 lanes): it reads the line's N values from LDS once and forms the N derivatives
// along it with the derivative matrix held in scalar registers.  The point-wise stage
// (strains, stress resultants, fluxes; force_excit_c.m:615-630) takes one point per lane.
// Derivatives and fluxes are stored line-major per direction (xi line j at j*N, eta line i
// at i*N, entry along the line last), so the transposed contractions work in place; the
// element forces are then xi part + eta part per point, plus the point terms of the two
// moments (Sxz W, Syz W).
//
// LDS: patch nodes [pn][5] (double), aliased by the derivative region [P][2 dir][5][NP];
// element values and forces [P][5][NP]; quadrature weight products [NP].
#include "sem_device.hpp"

namespace semk {

namespace {

// Derivative table, wavefront-uniform: double from the table, float from its float copy
// stored after the doubles (D, then the N weights).  The GLL derivative matrix is
// centro-antisymmetric, D(N-1-i, N-1-m) = -D(i, m), so only its first ceil(N/2) rows are
// read: half the scalar registers, the sign a free operand modifier.
template <typename T, int N>
struct ShellTab {
  const double* g;
  __device__ __forceinline__ T row(int i, int m) const {
    if constexpr (sizeof(T) == 8) return g[i * N + m];
    else return reinterpret_cast<const float*>(g + N * N + N)[i * N + m];
  }
  __device__ __forceinline__ T D(int i, int m) const {
    return 2 * i < N ? row(i, m) : -row(N - 1 - i, N - 1 - m);
  }
};

}  // namespace

// Whether the gathered patch-node rows stay intact beside the derivative region (FP32 form:
// its workgroups are limited by wavefronts, not LDS, so the rows cost no occupancy, and owned
// nodes take U^n from them at the update instead of reading it again); FP64 aliases them.
__host__ __device__ constexpr bool shell_keep_rows(bool fp32) { return fp32; }

__host__ __device__ constexpr size_t shell_rows_bytes(int max_pn) { return ((size_t)max_pn * 5 * 8 + 15) / 16 * 16; }

size_t shell_line_lds_bytes(int n, bool fp32, int max_pn, int P) {
  const size_t ts = fp32 ? 4 : 8, np = (size_t)n * n;
  const size_t r2 = shell_keep_rows(fp32) ? shell_rows_bytes(max_pn) + (ts * (size_t)P * 10 * np + 15) / 16 * 16
                                          : region2_bytes(max_pn, 5, ts * (size_t)P * 10 * np);
  return r2 + ts * ((size_t)P * 5 * np + np) + 16;
}

// FP32: 8 wavefronts per SIMD (two 16-element or four 8-element patches per CU)
template <int N, typename T, bool AFF, bool AK, bool GR, int P>
__global__ __launch_bounds__(64 * P, sizeof(T) == 4 ? 8 : 1) void shell_line_kernel(const double* __restrict__ gtab,
                                                                                    PatchArgs a, UpdateArgs ua) {
  constexpr int NP = N * N, NF = 5, TH = 64 * P, NT = 2 * NF * N;
  static_assert(NP <= 64, "one point per lane");
  extern __shared__ double smem[];
  constexpr bool KEEP = shell_keep_rows(sizeof(T) == 4);
  double* Up = smem;
  // [P][2 dir][5 field][NP], after the rows (KEEP) or over them
  T* dv = KEEP ? (T*)((char*)smem + shell_rows_bytes(a.max_pn)) : (T*)smem;
  const size_t dv_bytes = sizeof(T) * (size_t)P * 2 * NF * NP;
  T* su = (T*)((char*)smem + (KEEP ? shell_rows_bytes(a.max_pn) + (dv_bytes + 15) / 16 * 16
                                   : region2_bytes(a.max_pn, NF, dv_bytes)));  // [P][5][NP]
  T* sW = su + P * NF * NP;  // [NP] w_i w_j
  const ShellTab<T, N> tb{gtab};
  const int tid = threadIdx.x, b = xcd_block(blockIdx.x, gridDim.x);
  const int pn0 = a.pn_base[b], pn = a.pn_base[b + 1] - pn0;
  const int e0 = a.el0[b], ne = a.el0[b + 1] - e0;
  gather_patch<NF, TH, double, gather_per_thread(P * NP, TH)>(a, ua, Up, b, pn0, pn);
  OccHeld<1> occ;  // assembly indices of this thread's first patch node
  if constexpr (!GR) occ_prefetch<TH, 1>(a, b, pn0, pn, occ);
  if constexpr (AFF) {
    for (int t = tid; t < NP; t += TH) sW[t] = (T)(gtab[N * N + t % N] * gtab[N * N + t / N]);
  }
  __syncthreads();
  const int le = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int lane = tid & 63;
  const int64_t e = e0 + le;
  T* sue = su + le * NF * NP;
  T* dve = dv + le * 2 * NF * NP;
  if (le < ne && lane < NP) {
    const int ln = a.lmap[a.lm_base[b] + le * NP + lane];
#pragma unroll
    for (int f = 0; f < NF; ++f) sue[f * NP + lane] = (T)Up[ln * NF + f];
  }
  // the derivatives below overwrite the gathered patch nodes, unless they are kept (then each
  // wavefront only needs its own element's values)
  if constexpr (KEEP) wave_sync();
  else __syncthreads();
  // point p = i + N j of this lane; its entries in the xi and eta blocks
  const int p = lane, pi = p % N, pj = p / N;
  const int ix = p, ie = pi * N + pj;
  if (le < ne) {
    // --- forward: task t = (d*5 + f)*N + r; xi line r = j (values j*N + m), eta line r = i
    // (values m*N + i); the N derivatives go to dve[t*N ..]
    for (int t = lane; t < NT; t += 64) {
      const int df = t / N, r = t - df * N, d = df / NF, f = df - d * NF;
      const T* src = sue + f * NP + (d == 0 ? r * N : r);
      const int stride = d == 0 ? 1 : N;
      T u[N];
#pragma unroll
      for (int m = 0; m < N; ++m) u[m] = src[m * stride];
      T* dst = dve + t * N;
#pragma unroll
      for (int i = 0; i < N; ++i) {
        T s = 0;
#pragma unroll
        for (int m = 0; m < N; ++m) s += tb.D(i, m) * u[m];
        dst[i] = s;
      }
    }
    wave_sync();
    // --- point-wise: strains, stress resultants, fluxes in place
    T dmx = 0, dmy = 0;
    if (lane < NP) {
      T dxi[NF], deta[NF];
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        dxi[f] = dve[f * NP + ix];
        deta[f] = dve[(NF + f) * NP + ie];
      }
      const T phx = sue[1 * NP + p], phy = sue[3 * NP + p];
      T xx, xy, yx, yy, W;  // xi_x, xi_y, eta_x, eta_y (reference invJ11, invJ12, invJ21, invJ22)
      if constexpr (AFF) {
        const double* g = a.geom + e * 5;
        xx = (T)g[0]; xy = (T)g[1]; yx = (T)g[2]; yy = (T)g[3];
        W = sW[p] * (T)g[4];
      } else {
        T g[5];
        point_geom<T, 5>(a, e * NP + p, g);
        xx = g[0]; xy = g[1]; yx = g[2]; yy = g[3];
        W = g[4];
      }
      // material row as loaded; the element's stiffness scale multiplies the eight resultants
      // instead of the 21 coefficients
      const double* m = a.mat + (int64_t)a.elem_mat[e] * a.n_coef;
      const T sc = (a.sscale && !GR) ? (T)a.sscale[e] : (T)1;
      const T a11 = (T)m[0], a12 = (T)m[1], a16 = (T)m[2], a22 = (T)m[3], a26 = (T)m[4], a66 = (T)m[5],
              a44 = (T)m[6], a45 = (T)m[7], a55 = (T)m[8], b11 = (T)m[9], b12 = (T)m[10], b16 = (T)m[11],
              b22 = (T)m[12], b26 = (T)m[13], b66 = (T)m[14], d11 = (T)m[15], d12 = (T)m[16], d16 = (T)m[17],
              d22 = (T)m[18], d26 = (T)m[19], d66 = (T)m[20];
      // physical derivatives d/dx = d/dxi * xi_x + d/deta * eta_x (force_excit_c.m:615-622)
      const T u_x = dxi[0] * xx + deta[0] * yx, u_y = dxi[0] * xy + deta[0] * yy;
      const T fx_x = dxi[1] * xx + deta[1] * yx, fx_y = dxi[1] * xy + deta[1] * yy;
      const T v_x = dxi[2] * xx + deta[2] * yx, v_y = dxi[2] * xy + deta[2] * yy;
      const T fy_x = dxi[3] * xx + deta[3] * yx, fy_y = dxi[3] * xy + deta[3] * yy;
      const T w_x = dxi[4] * xx + deta[4] * yx, w_y = dxi[4] * xy + deta[4] * yy;
      const T ex = u_x, ey = v_y, gxy = u_y + v_x;
      const T kx = fx_x, ky = fy_y, kxy = fx_y + fy_x;
      const T gyz = phy + w_y, gxz = phx + w_x;  // '+' signs as coded in the reference
      const T Sxx = sc * (a11 * ex + a12 * ey + a16 * gxy + b11 * kx + b12 * ky + b16 * kxy);
      const T Syy = sc * (a12 * ex + a22 * ey + a26 * gxy + b12 * kx + b22 * ky + b26 * kxy);
      const T Sxy = sc * (a16 * ex + a26 * ey + a66 * gxy + b16 * kx + b26 * ky + b66 * kxy);
      const T Sxxz = sc * (b11 * ex + b12 * ey + b16 * gxy + d11 * kx + d12 * ky + d16 * kxy);
      const T Syyz = sc * (b12 * ex + b22 * ey + b26 * gxy + d12 * kx + d22 * ky + d26 * kxy);
      const T Sxyz = sc * (b16 * ex + b26 * ey + b66 * gxy + d16 * kx + d26 * ky + d66 * kxy);
      const T Syz = sc * (a44 * gyz + a45 * gxz);
      const T Sxz = sc * (a45 * gyz + a55 * gxz);
      // fluxes contracted by the transposed derivatives (force_excit_c.m:624-630)
      const T Wxx = W * xx, Wxy = W * xy, Wyx = W * yx, Wyy = W * yy;
      const T G0[NF] = {Sxx * Wxx + Sxy * Wxy, Sxxz * Wxx + Sxyz * Wxy, Sxy * Wxx + Syy * Wxy,
                        Sxyz * Wxx + Syyz * Wxy, Sxz * Wxx + Syz * Wxy};  // fu mx fv my fw, xi
      const T G1[NF] = {Sxx * Wyx + Sxy * Wyy, Sxxz * Wyx + Sxyz * Wyy, Sxy * Wyx + Syy * Wyy,
                        Sxyz * Wyx + Syyz * Wyy, Sxz * Wyx + Syz * Wyy};  // eta
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        dve[f * NP + ix] = G0[f];
        dve[(NF + f) * NP + ie] = G1[f];
      }
      dmx = Sxz * W;
      dmy = Syz * W;
    }
    wave_sync();
    // --- transposed contractions, every line in place
    for (int t = lane; t < NT; t += 64) {
      T* v = dve + t * N;
      T u[N];
#pragma unroll
      for (int m = 0; m < N; ++m) u[m] = v[m];
#pragma unroll
      for (int i = 0; i < N; ++i) {
        T s = 0;
#pragma unroll
        for (int m = 0; m < N; ++m) s += tb.D(m, i) * u[m];
        v[i] = s;
      }
    }
    wave_sync();
    T out[NF];
    if (lane < NP) {
#pragma unroll
      for (int f = 0; f < NF; ++f) out[f] = dve[f * NP + ix] + dve[(NF + f) * NP + ie];
      out[1] += dmx;
      out[3] += dmy;
      if (!AK && !GR && a.excite && a.ex_ptr) apply_excitation(a, ua, e, p, NF, out);
      if constexpr (!GR) {
#pragma unroll
        for (int f = 0; f < NF; ++f) sue[f * NP + p] = out[f];
      }
    }
    if constexpr (GR) {
      // per-point contributions to this element's gradient entries, summed by a fixed
      // shuffle tree over the wavefront
      double as = 0.0, am = 0.0;
      if (lane < NP) {
        const double* m = a.mat + (int64_t)a.elem_mat[e] * a.n_coef;
        const double Wp = AFF ? gtab[N * N + pi] * gtab[N * N + pj] * a.geom[e * 5 + 4]
                              : a.geom[(e * NP + p) * 5 + 4];
        const double mf[NF] = {m[21], m[22], m[21], m[22], m[21]};
        const int64_t nd = a.pn_id[pn0 + a.lmap[a.lm_base[b] + le * NP + p]];
#pragma unroll
        for (int f = 0; f < NF; ++f) {
          const int64_t d = nd * NF + f;
          const double lam = ua.lam[d], un = ua.Unew[d], u = ua.U[d], uo = ua.Uold[d];
          as += lam * (double)out[f];
          am += lam * Wp * mf[f] * ((un - 2.0 * u + uo) + ua.gdv[f] * (un - uo));
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
  patch_assemble<NF, NP, TH, AK, T, NF * NP, 1>(a, ua, su, b, pn0, pn, &occ, KEEP ? Up : nullptr);
}

namespace {

template <int N, typename T, bool AFF, bool AK, bool GR, int P>
hipError_t go(int64_t n_patch, const double* gtab, const PatchArgs& pa, const UpdateArgs& ua, hipStream_t s) {
  const size_t lds = shell_line_lds_bytes(N, sizeof(T) == 4, pa.max_pn, P);
  static LdsGrant granted;
  const hipError_t e = lds_opt_in(shell_line_kernel<N, T, AFF, AK, GR, P>, lds, granted);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((shell_line_kernel<N, T, AFF, AK, GR, P>), dim3((unsigned)n_patch), dim3(64 * P), lds, s, gtab,
                     pa, ua);
  return hipGetLastError();
}

template <int N, typename T, bool A, int P>
hipError_t by_mode(int mode, int64_t np, const double* g, const PatchArgs& pa, const UpdateArgs& ua, hipStream_t s) {
  if (mode == 0) return go<N, T, A, false, false, P>(np, g, pa, ua, s);
  if (mode == 1) return go<N, T, A, true, false, P>(np, g, pa, ua, s);
  return go<N, T, A, true, true, P>(np, g, pa, ua, s);
}

template <int N>
hipError_t by_type(bool fp32, bool aff, int mode, int P, int64_t np, const double* g, const PatchArgs& pa,
                   const UpdateArgs& ua, hipStream_t s) {
  if (P == 16) {
    if (fp32) return aff ? by_mode<N, float, true, 16>(mode, np, g, pa, ua, s)
                         : by_mode<N, float, false, 16>(mode, np, g, pa, ua, s);
    return aff ? by_mode<N, double, true, 16>(mode, np, g, pa, ua, s)
               : by_mode<N, double, false, 16>(mode, np, g, pa, ua, s);
  }
  if (P == 8) {
    if (fp32) return aff ? by_mode<N, float, true, 8>(mode, np, g, pa, ua, s)
                         : by_mode<N, float, false, 8>(mode, np, g, pa, ua, s);
    return aff ? by_mode<N, double, true, 8>(mode, np, g, pa, ua, s)
               : by_mode<N, double, false, 8>(mode, np, g, pa, ua, s);
  }
  return hipErrorInvalidValue;
}

}  // namespace

hipError_t launch_shell_line(int n, bool fp32, bool affine, int mode, int P, int64_t n_patch, const double* gtab,
                             const PatchArgs& pa, const UpdateArgs& ua, hipStream_t s) {
  switch (n) {
    case 3: return by_type<3>(fp32, affine, mode, P, n_patch, gtab, pa, ua, s);
    case 4: return by_type<4>(fp32, affine, mode, P, n_patch, gtab, pa, ua, s);
    case 5: return by_type<5>(fp32, affine, mode, P, n_patch, gtab, pa, ua, s);
    case 6: return by_type<6>(fp32, affine, mode, P, n_patch, gtab, pa, ua, s);
    case 7: return by_type<7>(fp32, affine, mode, P, n_patch, gtab, pa, ua, s);
    case 8: return by_type<8>(fp32, affine, mode, P, n_patch, gtab, pa, ua, s);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace semk
