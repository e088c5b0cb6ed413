// Flat-shell patch kernel with the GLL contractions on the matrix cores (gfx950, FP32 form).
//
// A workgroup holds one 16-element patch (576 threads, one per GLL point).  For the patch,
// both directional derivatives of all 5 fields are one product
//     dU [2 NP x 80] = Dk [2 NP x NP] . U [NP x 80]      (80 = 16 elements x 5 fields)
// with the Kronecker derivative matrix Dk = [D (x) I ; I (x) D] (rows: direction, point;
// columns: point), and the element forces of the transposed derivatives are
//     F [NP x 80] = Dk^T [NP x 2 NP] . G [2 NP x 80]      (G: the point fluxes).
// Both run as 16x16x16 bf16 MFMA tiles with FP32 accumulation.  Every FP32 operand x is
// split as x = hi + lo with hi = bf16(x), lo = bf16(x - hi) (16 significant bits), and a
// product is hi.hi + hi.lo + lo.hi, so a contraction carries a relative error near 2^-16
// against 2^-24 for FP32 arithmetic.  The point-wise stage (stress resultants, fluxes) is
// FP32 on the vector ALUs; gather, assembly and the update stay FP64 as in the other
// kernels.  Dk and Dk^T are precomputed on the host, split and laid out per lane (mtab).
#include <algorithm>

#include "sem_device.hpp"

namespace semk {

namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int ru16(int v) { return (v + 15) / 16 * 16; }

template <int N>
struct MShape {
  static constexpr int NP = N * N, P = 16, NF = 5, NC = P * NF;  // NC = 80 product columns
  static constexpr int M1 = ru16(2 * NP), K1 = ru16(NP);          // forward
  static constexpr int M2 = ru16(NP), K2 = ru16(2 * NP);          // transposed
  static constexpr int TH = P * NP;                               // one thread per point
  static constexpr int MW = TH / 64;  // full wavefronts: only they issue MFMA (no partial lanes)
  // LDS column strides (elements): the 16 columns one B-fragment load touches start on
  // different banks (8-byte loads at a stride of 2 x odd words: K is a multiple of 16, so
  // K + 4 bf16 is 8m + 2 words), and dU columns at an odd word stride
  static constexpr int LU = K1 + 4, LG = K2 + 4, LD1 = M1 + 1;
  static_assert(K1 % 16 == 0 && K2 % 16 == 0, "tiles of 16");
};

// round-to-nearest-even bf16 of a finite float, and its value
__device__ __forceinline__ unsigned short bf16_bits(float x) {
  const uint32_t b = __float_as_uint(x);
  return (unsigned short)((b + 0x7FFFu + ((b >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ float bf16_val(unsigned short h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ void split_store(unsigned short* hi, unsigned short* lo, int idx, float x) {
  const unsigned short h = bf16_bits(x);
  hi[idx] = h;
  lo[idx] = bf16_bits(x - bf16_val(h));
}

// C tile (rows mt*16.., columns nt*16..) += A . B over KT k-tiles; A fragments from the
// host table (per lane 4 bf16 = 8 bytes, hi then lo), B from LDS stored column-major
// (column n, k contiguous) as hi / lo arrays with leading dimension LD.
template <int KT, int LD>
__device__ __forceinline__ f32x4 tile_product(const uint64_t* __restrict__ A, int mt, const unsigned short* Bhi,
                                              const unsigned short* Blo, int nt, int lane) {
  f32x4 acc = {0.f, 0.f, 0.f, 0.f};
  const int n = nt * 16 + (lane & 15), k0 = 4 * (lane >> 4);
#pragma unroll
  for (int kt = 0; kt < KT; ++kt) {
    const uint64_t ah = A[((size_t)(mt * KT + kt) * 2 + 0) * 64 + lane];
    const uint64_t al = A[((size_t)(mt * KT + kt) * 2 + 1) * 64 + lane];
    const uint64_t bh = *reinterpret_cast<const uint64_t*>(Bhi + n * LD + kt * 16 + k0);
    const uint64_t bl = *reinterpret_cast<const uint64_t*>(Blo + n * LD + kt * 16 + k0);
    bf16x4 Ah, Al, Bh, Bl;
    __builtin_memcpy(&Ah, &ah, 8);
    __builtin_memcpy(&Al, &al, 8);
    __builtin_memcpy(&Bh, &bh, 8);
    __builtin_memcpy(&Bl, &bl, 8);
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(Ah, Bh, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(Ah, Bl, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(Al, Bh, acc, 0, 0, 0);
  }
  return acc;
}

}  // namespace

size_t shell_mfma_table_entries(int n);

size_t shell_mfma_lds_bytes(int n, int max_pn) {
  const int np = n * n, nc = 80, ld1 = ru16(2 * np) + 1, lu = ru16(np) + 4, lg = ru16(2 * np) + 4;
  const size_t regA = std::max({(size_t)max_pn * 5 * 8, (size_t)nc * ld1 * 4, (size_t)nc * lg * 2 * 2});
  const size_t regB = std::max((size_t)nc * lu * 2 * 2, (size_t)nc * np * 4);
  return (regA + 15) / 16 * 16 + (regB + 15) / 16 * 16 + shell_mfma_table_entries(n) * 8 + 16;
}

// Number of 8-byte entries of the A-operand table: forward then transposed tiles.
size_t shell_mfma_table_entries(int n) {
  const int np = n * n;
  return (size_t)(ru16(2 * np) / 16) * (ru16(np) / 16) * 2 * 64 + (size_t)(ru16(np) / 16) * (ru16(2 * np) / 16) * 2 * 64;
}

template <int N, bool AFF, bool AK, bool GR>
__global__ __launch_bounds__(MShape<N>::TH, 5) void shell_mfma_kernel(const double* __restrict__ gtab, PatchArgs a,
                                                                       UpdateArgs ua) {
  using S = MShape<N>;
  constexpr int NP = S::NP, NF = 5, TH = S::TH, NC = S::NC, M1 = S::M1, K1 = S::K1, M2 = S::M2, K2 = S::K2;
  constexpr int LU = S::LU, LG = S::LG, LD1 = S::LD1;
  extern __shared__ double smem[];
  // region A: gathered patch nodes (double), then dU [NC][LD1] (float), then G hi/lo [NC][LG]
  // region B: U hi/lo [NC][LU], then the element forces [NC][NP] (float)
  constexpr size_t regA0 = (size_t)NC * LD1 * 4 > (size_t)NC * LG * 4 ? (size_t)NC * LD1 * 4 : (size_t)NC * LG * 4;
  const size_t regA = ((regA0 > (size_t)a.max_pn * NF * 8 ? regA0 : (size_t)a.max_pn * NF * 8) + 15) / 16 * 16;
  double* Up = smem;
  float* dU = (float*)smem;
  unsigned short* Ghi = (unsigned short*)smem;
  unsigned short* Glo = Ghi + NC * LG;
  unsigned short* Uhi = (unsigned short*)((char*)smem + regA);
  unsigned short* Ulo = Uhi + NC * LU;
  float* Fo = (float*)((char*)smem + regA);
  // region C: the A-operand tables, staged once per patch (every tile reads them)
  constexpr size_t regB = (((size_t)NC * LU * 4 > (size_t)NC * NP * 4 ? (size_t)NC * LU * 4 : (size_t)NC * NP * 4) + 15) / 16 * 16;
  constexpr int NA = (M1 / 16) * (K1 / 16) * 2 * 64 + (M2 / 16) * (K2 / 16) * 2 * 64;
  uint64_t* sA = (uint64_t*)((char*)smem + regA + regB);
  for (int q = threadIdx.x; q < NA; q += TH) sA[q] = a.mtab[q];
  const uint64_t* Afw = sA;
  const uint64_t* Atr = sA + (size_t)(M1 / 16) * (K1 / 16) * 2 * 64;
  const int tid = threadIdx.x, b = xcd_block(blockIdx.x, gridDim.x), lane = tid & 63, wave = tid >> 6;
  const int pn0 = a.pn_base[b], pn = a.pn_base[b + 1] - pn0;
  const int e0 = a.el0[b], ne = a.el0[b + 1] - e0;
  gather_patch<NF, TH, double, gather_per_thread(S::P * NP, TH)>(a, ua, Up, b, pn0, pn);
  __syncthreads();
  const int le = tid / NP, p = tid - le * NP;
  const int i = p % N, j = p / N;
  const bool valid = le < ne;
  const int64_t e = e0 + le;
  float phx = 0.f, phy = 0.f;
  {  // B operand of the forward product: column (element, field), row = point
    float u[NF];
    if (valid) {
      const int ln = a.lmap[a.lm_base[b] + le * NP + p];
#pragma unroll
      for (int f = 0; f < NF; ++f) u[f] = (float)Up[ln * NF + f];
    } else {
#pragma unroll
      for (int f = 0; f < NF; ++f) u[f] = 0.f;
    }
    phx = u[1];
    phy = u[3];
#pragma unroll
    for (int f = 0; f < NF; ++f) split_store(Uhi, Ulo, (le * NF + f) * LU + p, u[f]);
    for (int q = tid; q < NC * (K1 - NP); q += TH) {  // zero padding rows
      const int c = q / (K1 - NP), r = NP + q % (K1 - NP);
      Uhi[c * LU + r] = 0;
      Ulo[c * LU + r] = 0;
    }
  }
  __syncthreads();  // dU overwrites the gathered nodes
  for (int t = wave; wave < S::MW && t < (M1 / 16) * (NC / 16); t += S::MW) {
    const int mt = t / (NC / 16), nt = t - mt * (NC / 16);
    const f32x4 c = tile_product<K1 / 16, LU>(Afw, mt, Uhi, Ulo, nt, lane);
    const int n = nt * 16 + (lane & 15);
#pragma unroll
    for (int r = 0; r < 4; ++r) dU[n * LD1 + mt * 16 + 4 * (lane >> 4) + r] = c[r];
  }
  __syncthreads();
  // point-wise stage (force_excit_c.m:615-630), one thread per point
  float dxi[NF], deta[NF];
#pragma unroll
  for (int f = 0; f < NF; ++f) {
    dxi[f] = dU[(le * NF + f) * LD1 + p];
    deta[f] = dU[(le * NF + f) * LD1 + NP + p];
  }
  __syncthreads();  // G overwrites dU
  float G[5][2];
  float dmx = 0.f, dmy = 0.f;
#pragma unroll
  for (int f = 0; f < 5; ++f) G[f][0] = G[f][1] = 0.f;
  if (valid) {
    float xx, xy, yx, yy, W;
    if (AFF) {
      const double* g = a.geom + e * 5;
      xx = (float)g[0]; xy = (float)g[1]; yx = (float)g[2]; yy = (float)g[3];
      W = (float)(gtab[N * N + i] * gtab[N * N + j]) * (float)g[4];
    } else {
      float g[5];
      point_geom<float, 5>(a, e * NP + p, g);
      xx = g[0]; xy = g[1]; yx = g[2]; yy = g[3];
      W = g[4];
    }
    const double* m = a.mat + (int64_t)a.elem_mat[e] * a.n_coef;
    const double sc = (a.sscale && !GR) ? a.sscale[e] : 1.0;
    const float a11 = (float)(m[0] * sc), a12 = (float)(m[1] * sc), a16 = (float)(m[2] * sc), a22 = (float)(m[3] * sc),
                a26 = (float)(m[4] * sc), a66 = (float)(m[5] * sc), a44 = (float)(m[6] * sc), a45 = (float)(m[7] * sc),
                a55 = (float)(m[8] * sc), b11 = (float)(m[9] * sc), b12 = (float)(m[10] * sc), b16 = (float)(m[11] * sc),
                b22 = (float)(m[12] * sc), b26 = (float)(m[13] * sc), b66 = (float)(m[14] * sc), d11 = (float)(m[15] * sc),
                d12 = (float)(m[16] * sc), d16 = (float)(m[17] * sc), d22 = (float)(m[18] * sc), d26 = (float)(m[19] * sc),
                d66 = (float)(m[20] * sc);
    const float u_x = dxi[0] * xx + deta[0] * yx, u_y = dxi[0] * xy + deta[0] * yy;
    const float fx_x = dxi[1] * xx + deta[1] * yx, fx_y = dxi[1] * xy + deta[1] * yy;
    const float v_x = dxi[2] * xx + deta[2] * yx, v_y = dxi[2] * xy + deta[2] * yy;
    const float fy_x = dxi[3] * xx + deta[3] * yx, fy_y = dxi[3] * xy + deta[3] * yy;
    const float w_x = dxi[4] * xx + deta[4] * yx, w_y = dxi[4] * xy + deta[4] * yy;
    const float ex = u_x, ey = v_y, gxy = u_y + v_x;
    const float kx = fx_x, ky = fy_y, kxy = fx_y + fy_x;
    const float gyz = phy + w_y, gxz = phx + w_x;  // '+' signs as coded in the reference
    const float Sxx = a11 * ex + a12 * ey + a16 * gxy + b11 * kx + b12 * ky + b16 * kxy;
    const float Syy = a12 * ex + a22 * ey + a26 * gxy + b12 * kx + b22 * ky + b26 * kxy;
    const float Sxy = a16 * ex + a26 * ey + a66 * gxy + b16 * kx + b26 * ky + b66 * kxy;
    const float Sxxz = b11 * ex + b12 * ey + b16 * gxy + d11 * kx + d12 * ky + d16 * kxy;
    const float Syyz = b12 * ex + b22 * ey + b26 * gxy + d12 * kx + d22 * ky + d26 * kxy;
    const float Sxyz = b16 * ex + b26 * ey + b66 * gxy + d16 * kx + d26 * ky + d66 * kxy;
    const float Syz = a44 * gyz + a45 * gxz;
    const float Sxz = a45 * gyz + a55 * gxz;
    const float Wxx = W * xx, Wxy = W * xy, Wyx = W * yx, Wyy = W * yy;
    G[0][0] = Sxx * Wxx + Sxy * Wxy;   G[0][1] = Sxx * Wyx + Sxy * Wyy;
    G[1][0] = Sxxz * Wxx + Sxyz * Wxy; G[1][1] = Sxxz * Wyx + Sxyz * Wyy;
    G[2][0] = Sxy * Wxx + Syy * Wxy;   G[2][1] = Sxy * Wyx + Syy * Wyy;
    G[3][0] = Sxyz * Wxx + Syyz * Wxy; G[3][1] = Sxyz * Wyx + Syyz * Wyy;
    G[4][0] = Sxz * Wxx + Syz * Wxy;   G[4][1] = Sxz * Wyx + Syz * Wyy;
    dmx = Sxz * W;
    dmy = Syz * W;
  }
#pragma unroll
  for (int f = 0; f < NF; ++f) {  // B operand of the transposed product: row = direction, point
    split_store(Ghi, Glo, (le * NF + f) * LG + p, G[f][0]);
    split_store(Ghi, Glo, (le * NF + f) * LG + NP + p, G[f][1]);
  }
  for (int q = tid; q < NC * (K2 - 2 * NP); q += TH) {
    const int c = q / (K2 - 2 * NP), r = 2 * NP + q % (K2 - 2 * NP);
    Ghi[c * LG + r] = 0;
    Glo[c * LG + r] = 0;
  }
  __syncthreads();
  for (int t = wave; wave < S::MW && t < (M2 / 16) * (NC / 16); t += S::MW) {
    const int mt = t / (NC / 16), nt = t - mt * (NC / 16);
    const f32x4 c = tile_product<K2 / 16, LG>(Atr, mt, Ghi, Glo, nt, lane);
    const int n = nt * 16 + (lane & 15);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int row = mt * 16 + 4 * (lane >> 4) + r;
      if (row < NP) Fo[n * NP + row] = c[r];
    }
  }
  __syncthreads();
  float* fe = Fo + le * NF * NP;
  if (valid) {
    fe[1 * NP + p] += dmx;
    fe[3 * NP + p] += dmy;
    if (!AK && !GR && a.excite && a.ex_ptr) {
      float out[NF];
#pragma unroll
      for (int f = 0; f < NF; ++f) out[f] = fe[f * NP + p];
      apply_excitation(a, ua, e, p, NF, out);
#pragma unroll
      for (int f = 0; f < NF; ++f) fe[f * NP + p] = out[f];
    }
  }
  if constexpr (GR) {
    __syncthreads();
    double* red = smem;  // the forces are read into registers first
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
        as += lam * (double)fe[f * NP + p];
        am += lam * Wp * mf[f] * ((un - 2.0 * u + uo) + ua.gdv[f] * (un - uo));
      }
    }
    __syncthreads();
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
  patch_assemble<NF, NP, TH, AK, float>(a, ua, Fo, b, pn0, pn);
}

namespace {

template <int N, bool AFF, bool AK, bool GR>
hipError_t go(int64_t n_patch, const double* gtab, const PatchArgs& pa, const UpdateArgs& ua, hipStream_t s) {
  const size_t lds = shell_mfma_lds_bytes(N, pa.max_pn);
  static LdsGrant granted;
  const hipError_t e = lds_opt_in(shell_mfma_kernel<N, AFF, AK, GR>, lds, granted);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((shell_mfma_kernel<N, AFF, AK, GR>), dim3((unsigned)n_patch), dim3(MShape<N>::TH), lds, s, gtab,
                     pa, ua);
  return hipGetLastError();
}

template <int N>
hipError_t by_mode(bool aff, int mode, int64_t np, const double* g, const PatchArgs& pa, const UpdateArgs& ua,
                   hipStream_t s) {
  if (aff) {
    if (mode == 0) return go<N, true, false, false>(np, g, pa, ua, s);
    if (mode == 1) return go<N, true, true, false>(np, g, pa, ua, s);
    return go<N, true, true, true>(np, g, pa, ua, s);
  }
  if (mode == 0) return go<N, false, false, false>(np, g, pa, ua, s);
  if (mode == 1) return go<N, false, true, false>(np, g, pa, ua, s);
  return go<N, false, true, true>(np, g, pa, ua, s);
}

}  // namespace

bool shell_mfma_supported(int n) { return n == 5 || n == 6; }

hipError_t launch_shell_mfma(int n, bool affine, int mode, int64_t n_patch, const double* gtab, const PatchArgs& pa,
                             const UpdateArgs& ua, hipStream_t s) {
  if (!pa.mtab) return hipErrorInvalidValue;
  if (n == 6) return by_mode<6>(affine, mode, n_patch, gtab, pa, ua, s);
  if (n == 5) return by_mode<5>(affine, mode, n_patch, gtab, pa, ua, s);
  return hipErrorInvalidValue;
}

}  // namespace semk
