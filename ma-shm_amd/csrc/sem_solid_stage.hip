// Hexahedral solid patch kernel with direction-staged GLL contractions (gfx950), variant 3.
//
// One wavefront per element, P elements (one patch) per workgroup, lanes own GLL lines as in
// the line-wise kernel (variant 1); what changes is what stays in LDS.  The line-wise kernel
// keeps all nine (direction, field) blocks of derivatives of an element in LDS at once, and
// at order 5 in FP64 that LDS limits a CU to two wavefronts per SIMD.  Here an element holds
// two [3][NP] blocks: its values V and one block B.  The derivatives of one direction go to
// B and are folded into per-point gradient registers (each lane keeps the points
// lane + 64 q), one direction after the other; stresses follow in registers; the fluxes of
// one direction go to B, are contracted in place with the transposed derivative matrix and
// added to per-point force registers, again one direction after the other.  The forces end
// in V (dead by then), where the patch assembly reads them.  6 NP values per element instead
// of 9 NP: at order 5 three FP32 8-element patches or three FP64 4-element patches share a CU
// (against two), at the price of per-direction passes over the points.
//
// Blocks are in point order p = i + N j + N^2 k, so the eta and zeta lines of a pass are
// read with unit stride across lanes.  The strains are accumulated direction by direction
// (six registers per point, not the nine gradient entries); forces sum (xi + eta) + zeta per
// point as in the line-wise kernel.
#include "sem_device.hpp"

namespace semk {

namespace {

// Lines of direction DIR (0 xi, 1 eta, 2 zeta) in a [3][NP] point-order block: line r of a
// field starts at base(r) and steps by STRIDE; xi lines are r = k N + j, eta lines
// r = k N + i, zeta lines r = j N + i.
template <int N, int NZ, int DIR>
struct Lines {
  static constexpr int NL = DIR == 2 ? N * N : N * NZ;  // lines per field
  static constexpr int LEN = DIR == 2 ? NZ : N;
  static constexpr int STRIDE = DIR == 0 ? 1 : (DIR == 1 ? N : N * N);
  __device__ __forceinline__ static int base(int r) {
    if constexpr (DIR == 0) return r * N;
    else if constexpr (DIR == 1) {
      const int k = r / N, i = r - k * N;
      return k * N * N + i;
    } else {
      return r;
    }
  }
};

// One pass over the 3 * NL lines of direction DIR: out = D u (TR = false) or D^T u (TR =
// true) along each line, from ld(field, point) into dst (in place when ld reads dst: a lane
// loads its whole line before it stores, and the lines of a pass are disjoint).
// The contraction is split into even and odd parts (the derivative matrix and its transpose
// are centro-antisymmetric): with e[m] = u[m] + u[L-1-m] and o[m] = u[m] - u[L-1-m], the
// outputs i and L-1-i are P + Q and P - Q of two half-length sums, 30 instead of 36 operations
// per line at L = 6 (coefficients from the even-odd tables, wavefront-uniform).
template <int N, int NZ, int DIR, bool TR, typename T, typename LD>
__device__ __forceinline__ void line_pass(const DerivTab<T, N, NZ>& tb, const LD& ld, T* dst, int lane) {
  using L = Lines<N, NZ, DIR>;
  constexpr int NP = N * N * NZ, LEN = L::LEN, H = LEN / 2;
  const T* A = tb.template eo<(DIR == 2 ? 2 : 0) + (TR ? 1 : 0)>();
  const T* B = A + H * H;
  const T* Cm = B + H * H;  // odd LEN: middle column, then middle row
  for (int t = lane; t < 3 * L::NL; t += 64) {
    const int f = t / L::NL, r = t - f * L::NL;
    const int q0 = L::base(r), o = f * NP + q0;
    T u[LEN];
#pragma unroll
    for (int m = 0; m < LEN; ++m) u[m] = ld(f, q0 + m * L::STRIDE);
    T ev[H], od[H];
#pragma unroll
    for (int m = 0; m < H; ++m) {
      ev[m] = u[m] + u[LEN - 1 - m];
      od[m] = u[m] - u[LEN - 1 - m];
    }
#pragma unroll
    for (int i = 0; i < H; ++i) {
      T P = 0, Q = 0;
#pragma unroll
      for (int m = 0; m < H; ++m) {
        P += A[i * H + m] * od[m];
        Q += B[i * H + m] * ev[m];
      }
      if constexpr (LEN % 2) Q += Cm[i] * u[H];
      dst[o + i * L::STRIDE] = P + Q;
      dst[o + (LEN - 1 - i) * L::STRIDE] = P - Q;
    }
    if constexpr (LEN % 2) {
      T s = 0;
#pragma unroll
      for (int m = 0; m < H; ++m) s += Cm[H + m] * od[m];
      dst[o + H * L::STRIDE] = s;
    }
  }
}

}  // namespace

// LDS layout.  FP32: [P][V | B] (the gathered patch nodes alias it) | weights [NP].
// FP64 ("rows"): the gathered patch-node rows | point -> patch node map [P][NP] (uint16) |
// [P][B] | 1D weights [N + NZ]; the element values are read through the map from the rows,
// so an element holds one [3][NP] block instead of two, and four 4-element patches of order
// 5 share a CU.
__host__ __device__ constexpr size_t stage_rows_bytes(int max_pn, size_t ts) { return ((size_t)max_pn * 3 * ts + 15) / 16 * 16; }
__host__ __device__ constexpr size_t stage_map_bytes(int P, int np) { return ((size_t)P * np * 2 + 15) / 16 * 16; }

// Whether a form reads its element values from the patch rows (rows in the arithmetic type):
// FP64.  In FP32 the rows would let four 8-element patches share a CU only at 64 VGPRs, where
// the kernel spills (72 bytes per lane); at 80 VGPRs three share one either way.
__host__ __device__ constexpr bool stage_rows(bool fp32) { return !fp32; }

size_t solid_stage_lds_bytes(int n, int nz, bool fp32, int max_pn, int P) {
  const size_t ts = fp32 ? 4 : 8, np = (size_t)n * n * nz;
  if (stage_rows(fp32)) return stage_rows_bytes(max_pn, ts) + stage_map_bytes(P, (int)np) + ts * ((size_t)P * 3 * np + n + nz) + 16;
  return region2_bytes(max_pn, 3, ts * (size_t)P * 6 * np) + ts * np + 16;
}

// Register budgets matched to the LDS: FP32 at most 80 VGPRs (6 wavefronts per SIMD: three
// 8-element patches per CU), FP64 at most 128 (4 per SIMD: four 4-element or two 8-element
// patches per CU)
// GEO: 0 per-point geometry, 1 affine element (one inverse Jacobian), 2 affine with a
// diagonal inverse Jacobian (axis-aligned; each gradient and flux entry then takes one
// direction's derivative).
template <int N, int NZ, typename T, int GEO, bool AK, bool GR, int P>
__global__ __launch_bounds__(64 * P) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 6 : 4)))
void solid_stage_kernel(const double* __restrict__ gtab, PatchArgs a, UpdateArgs ua) {
  constexpr int NP = N * N * NZ, NF = 3, TH = 64 * P, KP = (NP + 63) / 64;
  constexpr bool AFF = GEO > 0, DG = GEO == 2;
  constexpr bool ROWS = stage_rows(sizeof(T) == 4);  // values read from the patch-node rows
  constexpr int ES = ROWS ? 3 * NP : 6 * NP;  // LDS values per element
  extern __shared__ double smem[];
  using TR = typename std::conditional<ROWS, T, double>::type;  // row type
  TR* Up = (TR*)smem;
  uint16_t* sL = (uint16_t*)((char*)smem + stage_rows_bytes(a.max_pn, sizeof(TR)));  // rows: [P][NP]
  T* blk = ROWS ? (T*)((char*)sL + stage_map_bytes(P, NP))
                : (T*)smem;  // [P][ES]: rows [B]; else [V | B], the gathered rows aliasing it
  // weights: rows [w (N) | wz (NZ)], else their products per point [NP]
  T* sW = ROWS ? blk + P * ES : (T*)((char*)smem + region2_bytes(a.max_pn, NF, sizeof(T) * (size_t)P * 6 * NP));
  const DerivTab<T, N, NZ> tb{gtab};
  const int tid = threadIdx.x, b = xcd_block(blockIdx.x, gridDim.x);
  const int pn0 = a.pn_base[b], pn = a.pn_base[b + 1] - pn0;
  const int e0 = a.el0[b], ne = a.el0[b + 1] - e0;
  const int le = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int lane = tid & 63;
  // Every global load of the set-up is issued before the first LDS store: U^n of the
  // patch's nodes (at most P NP, i.e. KP per thread), the occurrence records held for the
  // assembly (below), and the local map of the thread's own element points.
  store_update_consts<NF>(ua);
  GatherHeld<NF, KP> gh;
  gather_load<NF, TH, KP>(a, ua, b, pn0, pn, 0, gh);
  // occurrence records of the thread's first patch nodes, held through the element stage:
  // two in FP64; none in the FP32 form, whose 80-VGPR budget they would strain (its
  // assembly loads them with the node states)
  constexpr int KH = sizeof(T) == 8 ? 2 : 0;
  OccHeld<KH> occ;
  if constexpr (!GR) occ_prefetch<TH, KH>(a, b, pn0, pn, occ);
  uint16_t lmv[KP];  // patch-local node of point lane + 64 k of the thread's element
  {
    const uint16_t* lm = a.lmap + a.lm_base[b] + (le < ne ? le : 0) * NP;
#pragma unroll
    for (int k = 0; k < KP; ++k) lmv[k] = lm[lane + 64 * k < NP ? lane + 64 * k : 0];
  }
  gather_store<NF, TH, KP>(gh, Up, pn, 0);
  if constexpr (AFF) {
    const double* w = gtab + N * N + NZ * NZ;
    if constexpr (ROWS) {
      for (int t = tid; t < N + NZ; t += TH) sW[t] = (T)w[t];
    } else {
      for (int t = tid; t < NP; t += TH) sW[t] = (T)(w[t % N] * w[(t / N) % N] * w[N + t / (N * N)]);
    }
  }
  // the map as byte offsets of the patch-node rows (at most P * NP nodes of NF values, within
  // 16 bits for the supported orders and patch sizes)
  static_assert(!ROWS || P * NP * NF * sizeof(TR) < 65536, "row offsets fit 16 bits");
  if constexpr (ROWS) {
    if (le < ne) {
#pragma unroll
      for (int k = 0; k < KP; ++k)
        if (lane + 64 * k < NP) sL[le * NP + lane + 64 * k] = (uint16_t)(lmv[k] * (NF * sizeof(TR)));
    }
  }
  __syncthreads();
  const int64_t e = e0 + le;
  T* V = blk + le * ES;  // !ROWS: the element values, later its forces
  T* B = ROWS ? V : V + 3 * NP;
  const uint16_t* Le = sL + le * NP;
  // the element's values by (field, point): from the rows through the map, or from V
  auto val = [&](int f, int q) -> T {
    if constexpr (ROWS) return (T) * (const TR*)((const char*)Up + Le[q] + f * sizeof(TR));
    else return V[f * NP + q];
  };
  auto blk_of = [&](T* X) { return [X](int f, int q) -> T { return X[f * NP + q]; }; };
  if constexpr (!ROWS) {
    T v[KP][NF];
    if (le < ne) {
#pragma unroll
      for (int k = 0; k < KP; ++k) {
        const int p = lane + 64 * k;
        if (p < NP) {
          const int ln = lmv[k];
#pragma unroll
          for (int f = 0; f < NF; ++f) v[k][f] = (T)Up[ln * NF + f];
        }
      }
    }
    __syncthreads();  // every patch-node read done before the element blocks are written
    if (le < ne) {
#pragma unroll
      for (int k = 0; k < KP; ++k) {
        const int p = lane + 64 * k;
        if (p < NP) {
#pragma unroll
          for (int f = 0; f < NF; ++f) V[f * NP + p] = v[k][f];
        }
      }
    }
  }
  if (le < ne) {
    wave_sync();
    const double* m = a.mat + (int64_t)a.elem_mat[e] * a.n_coef;
    const double sc = (a.sscale && !GR) ? a.sscale[e] : 1.0;
    T iJe[3][3], We = 0;
    if constexpr (AFF) {
      const double* g = a.geom + e * 10;
#pragma unroll
      for (int q = 0; q < 9; ++q) iJe[q / 3][q % 3] = (T)g[q];
      We = (T)g[9];
    }
    // row d of the inverse Jacobian at point p (and W = w_i w_j w_k detJ)
    auto geom_row = [&](int p, int d, T* r) {
      if constexpr (AFF) {
#pragma unroll
        for (int c = 0; c < 3; ++c) r[c] = iJe[d][c];
      } else {
        T g[10];
        point_geom<T, 10>(a, e * NP + p, g);
#pragma unroll
        for (int c = 0; c < 3; ++c) r[c] = g[3 * d + c];
      }
    };
    auto weight = [&](int p) -> T {
      if constexpr (AFF && ROWS) return sW[p % N] * sW[(p / N) % N] * sW[N + p / (N * N)] * We;
      else if constexpr (AFF) return sW[p] * We;
      else {
        T g[10];
        point_geom<T, 10>(a, e * NP + p, g);
        return g[9];
      }
    };
    // --- forward: strains Exx Eyy Ezz Exy Eyz Exz (strains_spec_p.m:8-27) from the gradient
    // d u_f / d x_c = sum_d (d u_f / d xi_d) iJ[d][c], accumulated one direction at a time (six
    // registers per point instead of the nine gradient entries)
    T E[KP][6];
    auto fold = [&](auto dirc) {
      constexpr int d = decltype(dirc)::value;
      wave_sync();
#pragma unroll
      for (int k = 0; k < KP; ++k) {
        const int p = lane + 64 * k;
        if (p < NP) {
          if constexpr (DG) {
            // this direction gives gradient column d only: E_dd, and one term of each shear with d
            T g[3];
#pragma unroll
            for (int f = 0; f < 3; ++f) g[f] = B[f * NP + p] * iJe[d][d];
            if constexpr (d == 0) {
              E[k][0] = g[0];
              E[k][3] = g[1];
              E[k][5] = g[2];
            } else if constexpr (d == 1) {
              E[k][1] = g[1];
              E[k][3] += g[0];
              E[k][4] = g[2];
            } else {
              E[k][2] = g[2];
              E[k][4] += g[1];
              E[k][5] += g[0];
            }
          } else {
            T r[3];
            geom_row(p, d, r);
            const T u0 = B[p], u1 = B[NP + p], u2 = B[2 * NP + p];
            if constexpr (d == 0) {
              E[k][0] = u0 * r[0];
              E[k][1] = u1 * r[1];
              E[k][2] = u2 * r[2];
              E[k][3] = u0 * r[1] + u1 * r[0];
              E[k][4] = u1 * r[2] + u2 * r[1];
              E[k][5] = u0 * r[2] + u2 * r[0];
            } else {  // chains of fused multiply-adds onto the running strains
              E[k][0] = E[k][0] + u0 * r[0];
              E[k][1] = E[k][1] + u1 * r[1];
              E[k][2] = E[k][2] + u2 * r[2];
              E[k][3] = E[k][3] + u0 * r[1] + u1 * r[0];
              E[k][4] = E[k][4] + u1 * r[2] + u2 * r[1];
              E[k][5] = E[k][5] + u0 * r[2] + u2 * r[0];
            }
          }
        }
      }
      wave_sync();  // the next direction overwrites B
    };
    line_pass<N, NZ, 0, false, T>(tb, val, B, lane);
    fold(std::integral_constant<int, 0>{});
    line_pass<N, NZ, 1, false, T>(tb, val, B, lane);
    fold(std::integral_constant<int, 1>{});
    line_pass<N, NZ, 2, false, T>(tb, val, B, lane);
    fold(std::integral_constant<int, 2>{});
    // --- stresses (stresses_spec_p.m:6-11), per point in registers
    const T D11 = (T)(m[0] * sc), D12 = (T)(m[1] * sc), D13 = (T)(m[2] * sc), D14 = (T)(m[3] * sc),
            D22 = (T)(m[4] * sc), D23 = (T)(m[5] * sc), D24 = (T)(m[6] * sc), D33 = (T)(m[7] * sc),
            D34 = (T)(m[8] * sc), D44 = (T)(m[9] * sc), D55 = (T)(m[10] * sc), D56 = (T)(m[11] * sc),
            D66 = (T)(m[12] * sc);
    T sg[KP][6];  // Sxx Syy Szz Sxy Syz Sxz
#pragma unroll
    for (int k = 0; k < KP; ++k) {
      const T Exx = E[k][0], Eyy = E[k][1], Ezz = E[k][2], Exy = E[k][3], Eyz = E[k][4], Exz = E[k][5];
      sg[k][0] = Exx * D11 + Eyy * D12 + Ezz * D13 + Exy * D14;
      sg[k][1] = Exx * D12 + Eyy * D22 + Ezz * D23 + Exy * D24;
      sg[k][2] = Exx * D13 + Eyy * D23 + Ezz * D33 + Exy * D34;
      sg[k][3] = Exx * D14 + Eyy * D24 + Ezz * D34 + Exy * D44;
      sg[k][4] = Eyz * D55 + Exz * D56;
      sg[k][5] = Eyz * D56 + Exz * D66;
    }
    // --- transposed: f_u += D_d^T (W sum_c sig[u][c] iJ[d][c]) (main_spec3Dcuda_global_..._mat.m:944-954),
    // one direction at a time, into per-point force registers
    T fe[KP][3] = {};
    auto flux_back = [&](auto dirc) {
      constexpr int d = decltype(dirc)::value;
#pragma unroll
      for (int k = 0; k < KP; ++k) {
        const int p = lane + 64 * k;
        if (p < NP) {
          if constexpr (d == 0) {
            // the point weight W = w_i w_j w_k detJ goes on the stresses once: the fluxes of
            // all three directions carry it
            const T W = weight(p);
#pragma unroll
            for (int c = 0; c < 6; ++c) sg[k][c] *= W;
          }
          const T Sxx = sg[k][0], Syy = sg[k][1], Szz = sg[k][2], Sxy = sg[k][3], Syz = sg[k][4], Sxz = sg[k][5];
          if constexpr (DG) {
            // W sigma[f][d]; the direction's inverse-Jacobian entry is applied to the
            // contracted fluxes
            const T sm[3][3] = {{Sxx, Sxy, Sxz}, {Sxy, Syy, Syz}, {Sxz, Syz, Szz}};
#pragma unroll
            for (int f = 0; f < 3; ++f) B[f * NP + p] = sm[f][d];
          } else {
            T r[3];
            geom_row(p, d, r);
            B[0 * NP + p] = Sxx * r[0] + Sxy * r[1] + Sxz * r[2];
            B[1 * NP + p] = Sxy * r[0] + Syy * r[1] + Syz * r[2];
            B[2 * NP + p] = Sxz * r[0] + Syz * r[1] + Szz * r[2];
          }
        }
      }
      wave_sync();
      line_pass<N, NZ, d, true, T>(tb, blk_of(B), B, lane);
      wave_sync();
#pragma unroll
      for (int k = 0; k < KP; ++k) {
        const int p = lane + 64 * k;
        if (p < NP) {
#pragma unroll
          for (int f = 0; f < 3; ++f) {
            const T x = B[f * NP + p];
            if constexpr (DG) fe[k][f] = d == 0 ? iJe[d][d] * x : fe[k][f] + iJe[d][d] * x;
            else fe[k][f] = d == 0 ? x : fe[k][f] + x;
          }
        }
      }
      wave_sync();  // the next direction overwrites B
    };
    flux_back(std::integral_constant<int, 0>{});
    flux_back(std::integral_constant<int, 1>{});
    flux_back(std::integral_constant<int, 2>{});
    // element forces into V (its values are no longer needed; rows: V is B)
#pragma unroll
    for (int k = 0; k < KP; ++k) {
      const int p = lane + 64 * k;
      if (p < NP) {
#pragma unroll
        for (int f = 0; f < 3; ++f) V[f * NP + p] = fe[k][f];
      }
    }
    wave_sync();
    // --- point sources (serial per element: entries may share a point)
    if constexpr (!AK && !GR) {
      if (a.excite && a.ex_ptr && lane == 0) {
        const int b0 = a.ex_ptr[e], b1 = a.ex_ptr[e + 1];
        if (b0 < b1) {
          const int64_t col = (int64_t)(ua.ctr[0] + ua.step_off) - 1;
          for (int q = b0; q < b1; ++q) {
            const int pf = a.ex_pf[q];
            V[(pf % NF) * NP + pf / NF] -= (T)(a.ex_amp[q] * a.st[(int64_t)a.ex_sig[q] * a.ex_nft + col]);
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
          const int64_t dd = nd * NF + f;
          const double lam = ua.lam[dd], un = ua.Unew[dd], u = ua.U[dd], uo = ua.Uold[dd];
          as += lam * (double)V[f * NP + p];
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
  // FP64 rows hold U^n exactly and are intact here: owned nodes take it from them
  const double* urow = nullptr;
  if constexpr (ROWS && std::is_same<TR, double>::value) urow = Up;
  // a thread's patch nodes are assembled two at a time, their record and state loads issued
  // together
  patch_assemble<NF, NP, TH, AK, T, ES, KH, 2>(a, ua, blk, b, pn0, pn, &occ, urow);
}

namespace {

template <int N, int NZ, typename T, int GEO, bool AK, bool GR, int P>
hipError_t go(int64_t n_patch, const double* gtab, const PatchArgs& pa, const UpdateArgs& ua, hipStream_t s) {
  const size_t lds = solid_stage_lds_bytes(N, NZ, sizeof(T) == 4, pa.max_pn, P);
  static LdsGrant granted;
  hipError_t e = lds_opt_in(solid_stage_kernel<N, NZ, T, GEO, AK, GR, P>, lds, granted);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((solid_stage_kernel<N, NZ, T, GEO, AK, GR, P>), dim3((unsigned)n_patch), dim3(64 * P), lds, s,
                     gtab, pa, ua);
  return hipGetLastError();
}

template <int N, int NZ, typename T, int GEO, int P>
hipError_t by_mode(int mode, int64_t np, const double* g, const PatchArgs& pa, const UpdateArgs& ua, hipStream_t s) {
  if (mode == 0) return go<N, NZ, T, GEO, false, false, P>(np, g, pa, ua, s);
  if (mode == 1) return go<N, NZ, T, GEO, true, false, P>(np, g, pa, ua, s);
  return go<N, NZ, T, GEO, true, true, P>(np, g, pa, ua, s);
}

template <int N, int NZ, typename T, int P>
hipError_t by_geo(bool aff, int mode, int64_t np, const double* g, const PatchArgs& pa, const UpdateArgs& ua,
                  hipStream_t s) {
  if (!aff) return by_mode<N, NZ, T, 0, P>(mode, np, g, pa, ua, s);
  // axis-aligned elements: the diagonal form saves a third of the point-wise products (both
  // forms; in FP64 it fits the 128-VGPR budget without spills since the point weight goes on
  // the stresses in the first flux direction)
  if (pa.diag) return by_mode<N, NZ, T, 2, P>(mode, np, g, pa, ua, s);
  return by_mode<N, NZ, T, 1, P>(mode, np, g, pa, ua, s);
}

template <int N, int NZ>
hipError_t by_type(bool fp32, bool aff, int mode, int P, int64_t np, const double* g, const PatchArgs& pa,
                   const UpdateArgs& ua, hipStream_t s) {
  if (P == 8) return fp32 ? by_geo<N, NZ, float, 8>(aff, mode, np, g, pa, ua, s)
                          : by_geo<N, NZ, double, 8>(aff, mode, np, g, pa, ua, s);
  if (P == 4) return fp32 ? by_geo<N, NZ, float, 4>(aff, mode, np, g, pa, ua, s)
                          : by_geo<N, NZ, double, 4>(aff, mode, np, g, pa, ua, s);
  return hipErrorInvalidValue;
}

}  // namespace

hipError_t launch_solid_stage(int n, int nz, bool fp32, bool affine, int mode, int P, int64_t n_patch,
                              const double* gtab, const PatchArgs& pa, const UpdateArgs& ua, hipStream_t s) {
#define STAGE_CASE(A, B) \
  if (n == A && nz == B) return by_type<A, B>(fp32, affine, mode, P, n_patch, gtab, pa, ua, s);
  STAGE_CASE(3, 3)
  STAGE_CASE(4, 4)
  STAGE_CASE(4, 3)
  STAGE_CASE(5, 5)
  STAGE_CASE(5, 3)
  STAGE_CASE(6, 6)
  STAGE_CASE(6, 3)
#undef STAGE_CASE
  return hipErrorInvalidValue;
}

}  // namespace semk
