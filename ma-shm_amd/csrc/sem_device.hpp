// Device helpers shared by the patch kernels (sem_kernels.hip, sem_solid_line.hip):
// launch shapes, LDS layout, wavefront exchange, node update, patch assembly, sources.
#pragma once
#include <atomic>
#include <type_traits>

#include "sem_host.hpp"
#include "sem_kernels.hpp"

namespace semk {

template <int N>
struct ShellP {
  static constexpr int NP = N * N;
  static constexpr int P = 16;
  static constexpr int TH = ((P * NP + 63) / 64) * 64;
};

template <int N, int NZ>
struct SolidP {
  static constexpr int NP = N * N * NZ;
  static constexpr int PPL = (NP + 63) / 64;
  static constexpr int P = 4;
  static constexpr int TH = 64 * P;
};


// Dynamic LDS above 64 KB needs an opt-in per kernel, which the runtime records per device.
// A launcher keeps one LdsGrant (a function-local static, so zero-initialised) with the
// largest size granted on each device; contexts on several threads may launch the same
// kernel, so the entries are atomic and a repeated opt-in is harmless.
struct LdsGrant {
  static constexpr int kMaxDevices = 64;
  std::atomic<size_t> size[kMaxDevices];
};

template <typename K>
inline hipError_t lds_opt_in(K kernel, size_t lds, LdsGrant& g) {
  if (lds <= 65536) return hipSuccess;
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  std::atomic<size_t>* granted = (dev >= 0 && dev < LdsGrant::kMaxDevices) ? &g.size[dev] : nullptr;
  if (granted && lds <= granted->load(std::memory_order_relaxed)) return hipSuccess;
  e = hipFuncSetAttribute(reinterpret_cast<const void*>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e == hipSuccess && granted) {
    size_t cur = granted->load(std::memory_order_relaxed);
    while (cur < lds && !granted->compare_exchange_weak(cur, lds, std::memory_order_relaxed)) {
    }
  }
  return e;
}

// LDS layout of a patch kernel: [R2 | su | tables].  R2 first holds the gathered patch
// nodes (double) and, once every element copied its points out of it, the fluxes.
__host__ __device__ constexpr size_t region2_bytes(int max_pn, int nf, size_t flux_bytes) {
  return ((((size_t)max_pn * nf * 8) > flux_bytes ? (size_t)max_pn * nf * 8 : flux_bytes) + 15) / 16 * 16;
}

// Workgroups are dispatched to the 8 XCDs in turn (workgroup i on XCD i % 8), and each XCD
// has its own L2.  This bijection over [0, n) gives XCD x one contiguous range of patches
// (XCDs below n % 8 get one more), so that neighbouring patches, which share boundary
// nodes, meet in the same L2.
__device__ __forceinline__ int xcd_block(int bid, int n) {
  const int x = bid & 7, k = bid >> 3, q = n >> 3, r = n & 7;
  return x * q + (x < r ? x : r) + k;
}

// LDS exchange inside one wavefront: DS instructions of a wavefront execute in order;
// the fences keep the compiler from moving accesses across this point.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ double coherent_load(const double* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ double sample(int q, double u, double uo, double un, double a0, double a1) {
  switch (q) {
    case 0: return u;
    case 1: return a1 * (-uo + un);
    case 2: return un;
    default: return a0 * (uo - 2.0 * u + un);
  }
}

__device__ __forceinline__ int frame_of(const UpdateArgs& ua, uint64_t nn) {
  return (int64_t)nn - 1 < ua.n_ftab ? ua.frame_tab[nn - 1] : -1;
}

// What the update of one node reads besides its force: U^n, U^{n-1} and the inverse
// lumped mass, all FP64 in both forms.  The inverse mass is stored per node, one column per
// distinct mass coefficient (mass_cols).
template <int NF, typename T>
struct NodeState {
  double u[NF], uo[NF];
  double mi[NF];
  __device__ __forceinline__ void load_u(const UpdateArgs& ua, int64_t node) {
#pragma unroll
    for (int f = 0; f < NF; ++f) u[f] = ua.U[node * NF + f];
  }
  __device__ __forceinline__ void load_rest(const UpdateArgs& ua, int64_t node) {
    constexpr int NM = mass_cols(NF);
    double m[NM];
#pragma unroll
    for (int c = 0; c < NM; ++c) m[c] = ua.minv[node * NM + c];
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      uo[f] = ua.Uold[node * NF + f];
      mi[f] = m[mass_col(NF, f)];
    }
  }
};

// Per-field constants c1, c2, c3 of the update, read from the kernel arguments ...
struct ArgConsts {
  const UpdateArgs& ua;
  __device__ __forceinline__ void at_use() const {}
  __device__ __forceinline__ double c1(int f) const { return ua.c1[f]; }
  __device__ __forceinline__ double c2(int f) const { return ua.c2[f]; }
  __device__ __forceinline__ double c3(int f) const { return ua.c3[f]; }
};

// ... or from a workgroup's LDS copy (patch kernels: held in scalar registers next to
// the patch kernels' other arguments they spill to vector lanes and are reloaded per use)
template <int NF>
__device__ __forceinline__ double* lds_consts() {
  __shared__ double s[3 * NF];
  return s;
}

template <int NF>
struct LdsConsts {
  const double* s;
  // read at each node rather than once per thread: a compiler barrier keeps the loads
  // next to their use instead of holding 3 x NF doubles in registers across the loop
  __device__ __forceinline__ void at_use() const { asm volatile("" ::: "memory"); }
  __device__ __forceinline__ double c1(int f) const { return s[f]; }
  __device__ __forceinline__ double c2(int f) const { return s[NF + f]; }
  __device__ __forceinline__ double c3(int f) const { return s[2 * NF + f]; }
};

// Central-difference update of one node (force_excit_c.m:687-693 rearranged as
// U^{n+1} = c1*minv*(-F) + c2*(2U^n - U^{n-1}) + c3*U^{n-1}, with per-field constants),
// plus frame output and divergence guard (force_excit_c.m:710-730).  The state
// arithmetic is FP64 in both forms.  Apply-K mode (AK) stores the force itself.
template <int NF, bool AK, typename T, typename C>
__device__ __forceinline__ void finish_node(const UpdateArgs& ua, int64_t node, const double* F,
                                            const NodeState<NF, T>& st, int fi, bool& bad, const C& cst) {
  const int64_t d0 = node * NF;
  if (AK) {
#pragma unroll
    for (int f = 0; f < NF; ++f) ua.Unew[d0 + f] = F[f];
    return;
  }
  double* fr = fi >= 0 ? ua.frames + (int64_t)fi * ua.dof : nullptr;
  cst.at_use();
#pragma unroll
  for (int f = 0; f < NF; ++f) {
    const double u = st.u[f], uo = st.uo[f];
    const double un = cst.c1(f) * st.mi[f] * (-F[f]) + cst.c2(f) * (2.0 * u - uo) + cst.c3(f) * uo;
    ua.Unew[d0 + f] = un;
    if (fi >= 0) {
      fr[d0 + f] = sample(ua.frame_q, u, uo, un, ua.a0, ua.a1);
      if (ua.guard_limit > 0.0 && (ua.guard_field < 0 || ua.guard_field == f)) {
        const double g = ua.guard_post ? un : u;
        if (g > ua.guard_limit) bad = true;
      }
      if (d0 + f == ua.nan_dof && un != un) bad = true;
    }
  }
}

// Internal id of a patch node.  The internal numbering hands out ids patch by patch in
// local order, owned nodes first, and an owned node occurs in one patch only, so a patch's
// owned nodes carry consecutive ids from pn_id[pn0] on; only boundary nodes read their id.
struct PatchNodes {
  const int32_t* id;
  int pn0, nown;
  int64_t own0;
  __device__ __forceinline__ PatchNodes(const PatchArgs& a, int b, int pn0_)
      : id(a.pn_id), pn0(pn0_), nown(a.pn_owned[b]), own0(a.pn_id[pn0_]) {}
  __device__ __forceinline__ int64_t operator()(int ln) const {
    return ln < nown ? own0 + ln : (int64_t)id[pn0 + ln];
  }
};

// Update constants for patch_assemble in the workgroup's LDS copy (read after a barrier).
template <int NF>
__device__ __forceinline__ void store_update_consts(const UpdateArgs& ua) {
  if (threadIdx.x == 0) {
    double* c = lds_consts<NF>();
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      c[f] = ua.c1[f];
      c[NF + f] = ua.c2[f];
      c[2 * NF + f] = ua.c3[f];
    }
  }
}

// U^n of KG patch nodes of a thread (node l0 + threadIdx.x + k TH), loaded in two batches:
// every boundary id first, then every value.  The loads carry no branches (a lane without a
// boundary node reads the patch's first id, a lane past the patch's nodes its first owned
// node, whose value it does not store), so that all of them are in flight together instead
// of one node's id and values after the other's.
template <int NF, int KG>
struct GatherHeld {
  double v[KG][NF];
};

template <int NF, int TH, int KG>
__device__ __forceinline__ void gather_load(const PatchArgs& a, const UpdateArgs& ua, int b, int pn0, int pn,
                                            int l0, GatherHeld<NF, KG>& g) {
  const int32_t* id = a.pn_id + pn0;
  const int nown = a.pn_owned[b];
  const int64_t own0 = id[0];
  int32_t x[KG];
#pragma unroll
  for (int k = 0; k < KG; ++k) {
    const int ln = l0 + (int)threadIdx.x + k * TH;
    x[k] = id[(ln >= nown && ln < pn) ? ln : 0];
  }
#pragma unroll
  for (int k = 0; k < KG; ++k) {
    const int ln = l0 + (int)threadIdx.x + k * TH;
    const int64_t nd = (ln >= nown && ln < pn) ? (int64_t)x[k] : own0 + (ln < pn ? ln : 0);
#pragma unroll
    for (int f = 0; f < NF; ++f) g.v[k][f] = ua.U[nd * NF + f];
  }
}

template <int NF, int TH, int KG, typename TR>
__device__ __forceinline__ void gather_store(const GatherHeld<NF, KG>& g, TR* Up, int pn, int l0) {
#pragma unroll
  for (int k = 0; k < KG; ++k) {
    const int ln = l0 + (int)threadIdx.x + k * TH;
    if (ln < pn) {
#pragma unroll
      for (int f = 0; f < NF; ++f) Up[ln * NF + f] = (TR)g.v[k][f];
    }
  }
}

// Patch nodes per thread that cover a patch of at most max_nodes nodes with TH threads.
__host__ __device__ constexpr int gather_per_thread(int max_nodes, int TH) { return (max_nodes + TH - 1) / TH; }

// Gather U^n of the patch's nodes into LDS rows [pn][NF] (double, or the arithmetic type
// of a kernel that reads its element values from them), KG nodes per thread at a time (a
// patch of P elements has at most P NP nodes: gather_per_thread(P NP, TH) covers it in one).
template <int NF, int TH, typename TR = double, int KG = 2>
__device__ __forceinline__ void gather_patch(const PatchArgs& a, const UpdateArgs& ua, TR* Up, int b,
                                             int pn0, int pn) {
  store_update_consts<NF>(ua);
  for (int l0 = 0; l0 < pn; l0 += KG * TH) {
    GatherHeld<NF, KG> g;
    gather_load<NF, TH, KG>(a, ua, b, pn0, pn, l0, g);
    gather_store<NF, TH, KG>(g, Up, pn, l0);
  }
}

// Occurrence record of a patch node (semhost::PatchPlan::occ8): its first kOccRec
// occurrences (LDS point indices le*NP + p) in list order, 16 bytes read with one load.
struct OccRec {
  uint32_t w[4];
  __device__ __forceinline__ int at(int j) const { return (int)((w[j >> 1] >> (16 * (j & 1))) & 0xFFFFu); }
};
static_assert(semhost::kOccRec == 8, "an occurrence record is 8 uint16 entries");

__device__ __forceinline__ OccRec load_occ_rec(const uint16_t* o8, int ln) {
  const uint4 r = *reinterpret_cast<const uint4*>(o8 + (size_t)ln * semhost::kOccRec);
  return OccRec{{r.x, r.y, r.z, r.w}};
}

// Occurrence records of the patch nodes a thread assembles in its first KH passes (node
// threadIdx.x + k*TH), loaded at the start of the patch kernel so that their latency
// overlaps the element stage instead of heading the assembly.  KH = 0 holds nothing.
template <int KH>
struct OccHeld {
  OccRec r[KH > 0 ? KH : 1];
};

template <int TH, int KH>
__device__ __forceinline__ void occ_prefetch(const PatchArgs& a, int b, int pn0, int pn, OccHeld<KH>& h) {
  const uint16_t* o8 = a.occ8 + a.o8_base[b];
#pragma unroll
  for (int k = 0; k < KH; ++k) {
    const int ln = threadIdx.x + k * TH;
    h.r[k] = load_occ_rec(o8, ln < pn ? ln : 0);
  }
}

// Patch assembly: every patch node sums its occurrences from the element-local forces
// in LDS (basket order); owned nodes are updated, boundary nodes write a partial (in the
// arithmetic type T: the FP32 form stores float partials) to the patch's slots.  Then the
// sensor terms of owned nodes are sampled into the ring.  Element le's forces are
// su[le*ES + f*NP + p].  A kernel whose gathered rows [pn][NF] of U^n are still in LDS
// passes them as urow, and owned nodes take U^n from there instead of reading it again.
// A node's occurrences come from its record (one load; the first KH nodes of a thread
// from held), those past a full record (kOccMore) from its occurrence list.
// KB > 0 takes a thread's nodes KB at a time (KB >= KH): the batch's global loads
// (occurrence records, U^{n-1} and inverse mass of its owned nodes) are all issued before
// its LDS sums, so that their latencies overlap instead of chaining node by node; it holds
// KB node states in registers, which suits kernels whose element stage has released them.
// KB = 0 takes one node at a time.
template <int NF, int NP, int TH, bool AK, typename T, int ES = NF * NP, int KH = 0, int KB = 0>
__device__ __forceinline__ void patch_assemble(const PatchArgs& a, const UpdateArgs& ua, const T* su, int b,
                                               int pn0, int pn, const OccHeld<KH>* held = nullptr,
                                               const double* urow = nullptr) {
  static_assert(KB == 0 || KB >= KH, "a batch holds the prefetched nodes");
  const uint16_t* o8 = a.occ8 + a.o8_base[b];
  const uint16_t* optr = a.occ_ptr + a.optr_base[b];
  const uint16_t* oc = a.occ + a.occ_base[b];
  const PatchNodes node(a, b, pn0);
  const int nown = node.nown;
  const uint64_t nn = AK ? 0 : ua.ctr[0] + ua.step_off;
  const int fi = AK ? -1 : frame_of(ua, nn);
  bool bad = false;
  // the force of patch node ln: its occurrences summed in list order
  auto node_force = [&](int ln, const OccRec& r, double* F) {
#pragma unroll
    for (int f = 0; f < NF; ++f) F[f] = 0.0;
    auto add = [&](int q) {
      const int le = q / NP, p = q - le * NP;
      const T* src = su + le * ES + p;
#pragma unroll
      for (int f = 0; f < NF; ++f) F[f] += (double)src[f * NP];
    };
    int q = 0;
#pragma unroll
    for (int j = 0; j < semhost::kOccRec; ++j) {
      q = r.at(j);
      if (q >= semhost::kOccMore) break;
      add(q);
    }
    if (q == semhost::kOccMore)
      for (int o = optr[ln] + semhost::kOccRec - 1; o < optr[ln + 1]; ++o) add(oc[o]);
  };
  auto finish = [&](int ln, const double* F, const NodeState<NF, T>& st) {
    if (ln < nown) {
      finish_node<NF, AK, T>(ua, node(ln), F, st, fi, bad, LdsConsts<NF>{lds_consts<NF>()});
    } else {
      T* dst = (T*)ua.partial + (int64_t)(a.slot_base[b] + ln - nown) * NF;
#pragma unroll
      for (int f = 0; f < NF; ++f) dst[f] = (T)F[f];
    }
  };
  auto load_state = [&](int ln, NodeState<NF, T>& st) {
    if (!AK && ln < nown) {
      const int64_t nd = node(ln);
      if (urow) {
#pragma unroll
        for (int f = 0; f < NF; ++f) st.u[f] = urow[ln * NF + f];
      } else {
        st.load_u(ua, nd);
      }
      st.load_rest(ua, nd);
    }
  };
  if constexpr (KB == 0) {
    auto node_step = [&](int ln, const OccRec& r) {
      double F[NF];
      node_force(ln, r, F);
      NodeState<NF, T> st;
      load_state(ln, st);
      finish(ln, F, st);
    };
#pragma unroll
    for (int k = 0; k < KH; ++k) {
      const int ln = threadIdx.x + k * TH;
      if (ln < pn) node_step(ln, held->r[k]);
    }
    for (int ln = threadIdx.x + KH * TH; ln < pn; ln += TH) node_step(ln, load_occ_rec(o8, ln));
  } else {
    constexpr int KHB = KH > 0 ? KH : 1;
    for (int j0 = 0; threadIdx.x + j0 * TH < pn; j0 += KB) {
      const bool first = j0 == 0;  // the batch whose first KH nodes were prefetched
      OccRec r[KB];
      NodeState<NF, T> st[KB];
#pragma unroll
      for (int k = 0; k < KB; ++k) {
        const int ln = threadIdx.x + (j0 + k) * TH;
        if (k < KH && first) r[k] = held->r[k % KHB];
        else r[k] = load_occ_rec(o8, ln < pn ? ln : 0);
      }
#pragma unroll
      for (int k = 0; k < KB; ++k) {
        const int ln = threadIdx.x + (j0 + k) * TH;
        if (ln < pn) load_state(ln, st[k]);
      }
#pragma unroll
      for (int k = 0; k < KB; ++k) {
        const int ln = threadIdx.x + (j0 + k) * TH;
        if (ln >= pn) continue;
        double F[NF];
        node_force(ln, r[k], F);
        finish(ln, F, st[k]);
      }
    }
  }
  if (bad) atomicMin((unsigned long long*)&ua.ctr[1], (unsigned long long)nn);
  if (!AK && a.tp_ptr) {
    const int t0 = a.tp_ptr[b], t1 = a.tp_ptr[b + 1];
    if (t0 < t1) {
      __syncthreads();
      for (int q = t0 + threadIdx.x; q < t1; q += TH) {
        const int lf = a.tp_lf[q], ln = lf / NF, f = lf - ln * NF;
        const int64_t d = node(ln) * NF + f;
        ua.ring[ua.step_off * ua.n_terms + a.tp_t[q]] =
            sample(ua.out_q, ua.U[d], ua.Uold[d], coherent_load(ua.Unew + d), ua.a0, ua.a1);
      }
    }
  }
}

// Derivative tables, wavefront-uniform: double from the table, float from its float copy
// stored after the doubles (the kernels read the same values in the arithmetic type).
// The GLL derivative matrices are centro-antisymmetric, D(N-1-i, N-1-m) = -D(i, m), so only
// their first ceil(N/2) rows are read: half the scalar registers, the sign a free operand
// modifier.
template <typename T, int N, int NZ>
struct DerivTab {
  const double* g;
  __device__ __forceinline__ T row(int i, int m) const {
    if constexpr (sizeof(T) == 8) return g[i * N + m];
    else return reinterpret_cast<const float*>(g + N * N + NZ * NZ + N + NZ)[i * N + m];
  }
  __device__ __forceinline__ T rowz(int i, int m) const {
    if constexpr (sizeof(T) == 8) return g[N * N + i * NZ + m];
    else return reinterpret_cast<const float*>(g + N * N + NZ * NZ + N + NZ)[N * N + i * NZ + m];
  }
  __device__ __forceinline__ T D(int i, int m) const {
    return 2 * i < N ? row(i, m) : -row(N - 1 - i, N - 1 - m);
  }
  __device__ __forceinline__ T Dz(int i, int m) const {
    return 2 * i < NZ ? rowz(i, m) : -rowz(NZ - 1 - i, NZ - 1 - m);
  }
  // Even-odd tables (solids; semhost::even_odd_table), after the doubles and their float
  // copy: for Q = 0 D, 1 D^T, 2 Dz, 3 Dz^T of size L x L, A [h][h], B [h][h] and for odd L
  // the middle column [h] and middle row [h], h = L / 2; doubles first, then a float copy.
  static constexpr int eo_size(int L) { return 2 * (L / 2) * (L / 2) + (L % 2) * 2 * (L / 2); }
  static constexpr int EO_D = 2 * eo_size(N) + 2 * eo_size(NZ);
  static constexpr int ND = N * N + NZ * NZ + N + NZ;
  template <int Q>
  __device__ __forceinline__ const T* eo() const {
    constexpr int off = (Q >= 1 ? eo_size(N) : 0) + (Q >= 2 ? eo_size(N) : 0) + (Q >= 3 ? eo_size(NZ) : 0);
    const double* base = g + ND + (ND + 1) / 2;
    if constexpr (sizeof(T) == 8) return base + off;
    else return reinterpret_cast<const float*>(base + EO_D) + off;
  }
};

// Per-point geometry of a non-affine element in the arithmetic type (the FP32 form reads
// its float copy: half the bytes of the per-point stream).
template <typename T, int G>
__device__ __forceinline__ void point_geom(const PatchArgs& a, int64_t idx, T* out) {
  if constexpr (sizeof(T) == 4) {
    const float* g = a.geomf + idx * G;
#pragma unroll
    for (int q = 0; q < G; ++q) out[q] = g[q];
  } else {
    const double* g = a.geom + idx * G;
#pragma unroll
    for (int q = 0; q < G; ++q) out[q] = g[q];
  }
}

template <typename T>
__device__ __forceinline__ void apply_excitation(const PatchArgs& a, const UpdateArgs& ua, int64_t e, int p,
                                                 int nf, T* out) {
  const int b0 = a.ex_ptr[e], b1 = a.ex_ptr[e + 1];
  if (b0 < b1) {
    const int64_t col = (int64_t)(ua.ctr[0] + ua.step_off) - 1;
    for (int k = b0; k < b1; ++k) {
      const int pf = a.ex_pf[k];
      if (pf / nf == p) out[pf % nf] -= (T)(a.ex_amp[k] * a.st[(int64_t)a.ex_sig[k] * a.ex_nft + col]);
    }
  }
}

}  // namespace semk
