// C-ABI of the spectral-element library: context, uploads, the time loop.
// Declarations and the reference interfaces they replace: include/sem_hip.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/sem_hip.h"
#include "sem_host.hpp"
#include "sem_kernels.hpp"

// steps per captured graph (a multiple of 3 keeps the state-buffer rotation periodic)
constexpr int kChunk = 30;

// Point sources on the device, attached to element-local points in patch element order:
// entries ptr[e]..ptr[e+1] of element e add -amp * st[sig * nft + nn - 1] to field pf % nf
// of point pf / nf.
struct SourceSet {
  int32_t *ptr = nullptr, *pf = nullptr, *sig = nullptr;
  double *amp = nullptr, *st = nullptr;
  int64_t nft = 0;
  void release() {
    for (void* p : {(void*)ptr, (void*)pf, (void*)sig, (void*)amp, (void*)st})
      if (p) hipFree(p);
    ptr = pf = sig = nullptr;
    amp = st = nullptr;
    nft = 0;
  }
};

struct sem_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  // mesh
  int kind = 0, nf = 0, nxy = 0, nz = 1, nloc = 0;
  int64_t n_el = 0, n_nodes = 0, dof = 0;
  bool affine = true;
  bool diag = false;   // affine solid whose inverse Jacobians are diagonal (axis-aligned elements)
  int geom_stride = 0;
  std::vector<int32_t> h_nodes;
  std::vector<double> h_coords;
  std::vector<double> h_geom;      // per element or per point
  std::vector<double> h_wdetj;     // per point W*detJ (mass assembly)
  std::vector<int8_t> basket;
  std::vector<int32_t> brow;
  int64_t rows = 0;
  std::vector<int32_t> rowptr, occ;
  std::vector<double> h_mass;
  double dv[5] = {0, 0, 0, 0, 0};
  // device
  double* d_gtab = nullptr;
  double* d_geom = nullptr;
  float* d_geomf = nullptr;  // per-point geometry as float (non-affine meshes)
  semhost::PatchPlan plan;
  std::vector<int32_t> inv_perm;   // caller's element -> patch order
  int32_t* d_node_new = nullptr;   // caller's node -> internal node, on the device
  int64_t n_used = 0;              // nodes that belong to an element: internal ids 0..n_used-1
  double* scratch_dev = nullptr;   // device address of `scratch` on this context's device
  int32_t *d_el0 = nullptr, *d_pn_base = nullptr, *d_pn_owned = nullptr, *d_pn_id = nullptr,
          *d_slot_base = nullptr, *d_occ_base = nullptr, *d_sh_ptr = nullptr, *d_sh_slot = nullptr,
          *d_sh_id = nullptr, *d_optr_base = nullptr, *d_lm_base = nullptr, *d_o8_base = nullptr;
  uint16_t *d_occ_ptr = nullptr, *d_pocc = nullptr, *d_lmap = nullptr, *d_occ8 = nullptr;
  uint64_t* d_mtab = nullptr;  // MFMA shell kernel tables (shells of supported order)
  double* d_partial = nullptr;
  double* d_mat = nullptr;
  int n_coef = 0;
  int32_t* d_elem_mat = nullptr;
  double* d_sscale = nullptr;
  double* d_minv = nullptr;
  double* d_u[3] = {nullptr, nullptr, nullptr};
  int rot = 0;
  bool have_material = false;
  // excitation
  SourceSet ex;
  // outputs: sensor terms t = (dof, weight) grouped per output; each term is sampled by
  // the kernel that updates its node (owned: patch list, boundary: shared-node list)
  int n_out = 0, out_q = 0;
  int64_t n_terms = 0;
  int64_t* d_out_ptr = nullptr;
  double* d_out_w = nullptr;
  std::vector<int64_t> h_out_ptr, h_out_dof;
  std::vector<double> h_out_w;
  int32_t *d_tp_ptr = nullptr, *d_tp_lf = nullptr, *d_tp_t = nullptr;
  int32_t *d_ts_ptr = nullptr, *d_ts_f = nullptr, *d_ts_t = nullptr;
  double* d_ring = nullptr;
  double* d_traces = nullptr;
  int64_t traces_cap = 0;
  std::vector<int32_t> owner_patch, owner_ln, shared_idx;  // per node
  // frames
  int n_frames = 0, frame_q = 1;
  std::vector<int64_t> h_frame_nn;
  int64_t* d_frame_nn = nullptr;
  int32_t* d_frame_tab = nullptr;
  int64_t n_ftab = 0;
  double* d_frames = nullptr;
  // guard
  double guard_limit = 0.0;
  int guard_field = -1, guard_post = 0;
  uint64_t* d_ctr = nullptr;  // semk::kCtr counters: chunk's first step, guard step
  // element-kernel choice: variant, patch size requested for the next upload (0 = default)
  int variant = 0, patch_req = 0;
  // graph cache
  hipGraphExec_t graph = nullptr;
  int graph_rot = -1, graph_steps = 0, graph_prec = -1;
  double graph_dt = 0.0;
  int64_t graph_base = 0;
  int64_t graph_ld = 0;
};

// Scratch of the allocation-free entry point sem_apply_K_noalloc: the internal-order copies
// of u and f and the boundary nodes' partial sums.  Each device holds its own copy of this
// array from the first use of the library on it until the process ends.  In host code the
// array's name is not a device address: every host use goes through the pointer that
// hipGetSymbolAddress returned for the device (scratch_on).
constexpr size_t kScratchBytes = size_t(128) << 20;
static __device__ __align__(256) double scratch[kScratchBytes / sizeof(double)];

namespace {

constexpr int kMaxScratchDevices = 64;
std::atomic<double*> g_scratch[kMaxScratchDevices];  // per device, set once

// The device address of `scratch` on `device` (which must be current), looked up once per
// device.  Returns nullptr and sets *e when the lookup fails or the device id is too large.
double* scratch_on(int device, hipError_t* e) {
  *e = hipSuccess;
  if (device < 0 || device >= kMaxScratchDevices) {
    *e = hipErrorInvalidDevice;
    return nullptr;
  }
  double* p = g_scratch[device].load(std::memory_order_acquire);
  if (p) return p;
  *e = hipGetSymbolAddress((void**)&p, HIP_SYMBOL(scratch));
  if (*e != hipSuccess) return nullptr;
  g_scratch[device].store(p, std::memory_order_release);
  return p;
}

// [n_nodes x nf] between the caller's node order and the internal one, on the device.
// Out: nodes that belong to no element (internal ids from n_used on) give 0.
__global__ __launch_bounds__(256) void permute_in_kernel(const double* __restrict__ ext, double* __restrict__ in,
                                                         const int32_t* __restrict__ node_new, int64_t n, int nf) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * nf) return;
  const int64_t k = t / nf;
  in[(int64_t)node_new[k] * nf + (t - k * nf)] = ext[t];
}

__global__ __launch_bounds__(256) void permute_out_kernel(const double* __restrict__ in, double* __restrict__ ext,
                                                          const int32_t* __restrict__ node_new, int64_t n, int nf,
                                                          int64_t n_used) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * nf) return;
  const int64_t k = t / nf, nd = node_new[k];
  ext[t] = nd < n_used ? in[nd * nf + (t - k * nf)] : 0.0;
}

int fail(sem_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

#define HIPCHK(c, call)                                                                   \
  do {                                                                                    \
    hipError_t _e = (call);                                                               \
    if (_e != hipSuccess)                                                                 \
      return fail((c), SEM_E_HIP, std::string(#call) + ": " + hipGetErrorString(_e));    \
  } while (0)

// HIP events owned by one call: destroyed on every return path.
struct Events {
  std::vector<hipEvent_t> ev;
  hipError_t create(size_t n) {
    ev.reserve(n);
    for (size_t k = 0; k < n; ++k) {
      hipEvent_t e = nullptr;
      const hipError_t r = hipEventCreate(&e);
      if (r != hipSuccess) return r;
      ev.push_back(e);
    }
    return hipSuccess;
  }
  hipEvent_t& operator[](size_t k) { return ev[k]; }
  size_t size() const { return ev.size(); }
  ~Events() {
    for (hipEvent_t e : ev) hipEventDestroy(e);
  }
};

template <typename P>
void dfree(P*& p) {
  if (p) hipFree((void*)p);
  p = nullptr;
}

template <typename P>
hipError_t dupload(P*& dst, const P* src, size_t n, hipStream_t s) {
  dfree(dst);
  if (n == 0) return hipSuccess;
  hipError_t e = hipMalloc((void**)&dst, n * sizeof(P));
  if (e != hipSuccess) return e;
  e = hipMemcpyAsync(dst, src, n * sizeof(P), hipMemcpyHostToDevice, s);
  if (e != hipSuccess) return e;
  return hipStreamSynchronize(s);
}

void drop_graph(sem_ctx* c) {
  if (c->graph) hipGraphExecDestroy(c->graph);
  c->graph = nullptr;
  c->graph_rot = -1;
}

// Per-point geometry of one element (force_excit_c.m:328-346,406-416; solid :480-604).
// Shell output per point: invJ11 invJ12 invJ21 invJ22 detJ.
// Solid output per point: invJ (3x3 row-major, iJ[d][c] = d xi_d/d x_c), detJ.
void element_geometry(const sem_ctx* c, int64_t e, const std::vector<double>& D,
                      const std::vector<double>& Dz, std::vector<double>& out) {
  const int n = c->nxy, nz = c->nz, np = c->nloc;
  const int32_t* nd = c->h_nodes.data() + e * np;
  if (c->kind == SEM_SHELL5) {
    out.assign((size_t)np * 5, 0.0);
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i) {
        double xk = 0, xe = 0, yk = 0, ye = 0;
        for (int k = 0; k < n; ++k) {
          const double* pk = &c->h_coords[3 * (int64_t)nd[j * n + k]];
          const double* pe = &c->h_coords[3 * (int64_t)nd[k * n + i]];
          xk += D[i * n + k] * pk[0];
          yk += D[i * n + k] * pk[1];
          xe += D[j * n + k] * pe[0];
          ye += D[j * n + k] * pe[1];
        }
        const double J11 = xk, J21 = yk, J12 = xe, J22 = ye;  // J11=Npx*X, J21=Npx*Y, J12=Npy*X
        const double det = J11 * J22 - J12 * J21;
        const double a = 1.0 / det;
        double* o = &out[(size_t)(j * n + i) * 5];
        o[0] = a * J22;
        o[1] = a * (-J12);
        o[2] = a * (-J21);
        o[3] = a * J11;
        o[4] = det;
      }
  } else {
    out.assign((size_t)np * 10, 0.0);
    for (int k = 0; k < nz; ++k)
      for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
          double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};  // J[a][d] = d x_a / d xi_d
          for (int m = 0; m < n; ++m) {
            const double* p1 = &c->h_coords[3 * (int64_t)nd[(k * n + j) * n + m]];
            const double* p2 = &c->h_coords[3 * (int64_t)nd[(k * n + m) * n + i]];
            for (int a = 0; a < 3; ++a) {
              J[a][0] += D[i * n + m] * p1[a];
              J[a][1] += D[j * n + m] * p2[a];
            }
          }
          for (int m = 0; m < nz; ++m) {
            const double* p3 = &c->h_coords[3 * (int64_t)nd[(m * n + j) * n + i]];
            for (int a = 0; a < 3; ++a) J[a][2] += Dz[k * nz + m] * p3[a];
          }
          // inverse by the adjugate (the reference's inv_jacp/det_jacp are absent from
          // its tree; this is the 3x3 inverse they are called for, :544-545)
          const double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                             J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                             J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
          const double a = 1.0 / det;
          double inv[3][3];
          inv[0][0] = a * (J[1][1] * J[2][2] - J[1][2] * J[2][1]);
          inv[0][1] = a * (J[0][2] * J[2][1] - J[0][1] * J[2][2]);
          inv[0][2] = a * (J[0][1] * J[1][2] - J[0][2] * J[1][1]);
          inv[1][0] = a * (J[1][2] * J[2][0] - J[1][0] * J[2][2]);
          inv[1][1] = a * (J[0][0] * J[2][2] - J[0][2] * J[2][0]);
          inv[1][2] = a * (J[0][2] * J[1][0] - J[0][0] * J[1][2]);
          inv[2][0] = a * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
          inv[2][1] = a * (J[0][1] * J[2][0] - J[0][0] * J[2][1]);
          inv[2][2] = a * (J[0][0] * J[1][1] - J[0][1] * J[1][0]);
          double* o = &out[(size_t)((k * n + j) * n + i) * 10];
          for (int d = 0; d < 3; ++d)
            for (int cc = 0; cc < 3; ++cc) o[3 * d + cc] = inv[d][cc];
          o[9] = det;
        }
  }
}

// bf16 bits of a float, round to nearest even (finite values)
uint16_t bf16_bits(float x) {
  uint32_t b;
  std::memcpy(&b, &x, 4);
  return (uint16_t)((b + 0x7FFFu + ((b >> 16) & 1u)) >> 16);
}

float bf16_value(uint16_t h) {
  const uint32_t b = (uint32_t)h << 16;
  float x;
  std::memcpy(&x, &b, 4);
  return x;
}

// A operands of the MFMA shell kernel: the Kronecker derivative matrix Dk [2NP x NP]
// (row dir*NP + (j*n + i): d/dxi (dir 0) or d/deta (dir 1) at point (i, j); column: point)
// and its transpose, zero-padded to multiples of 16, split into bf16 hi + lo, and laid
// out per 16x16x16 tile as the lanes read them: tile (mt, kt), part (hi, lo), lane l holds
// rows mt*16 + l%16, columns kt*16 + 4*(l/16) + 0..3.
std::vector<uint64_t> mfma_tables(int n, const std::vector<double>& D) {
  const int np = n * n, r16 = 16;
  auto ru = [&](int v) { return (v + r16 - 1) / r16 * r16; };
  auto dk = [&](int row, int col) -> double {  // Dk
    if (row >= 2 * np || col >= np) return 0.0;
    const int dir = row / np, pp = row % np, i = pp % n, j = pp / n, kk = col % n, l = col / n;
    return dir == 0 ? (l == j ? D[i * n + kk] : 0.0) : (kk == i ? D[j * n + l] : 0.0);
  };
  std::vector<uint64_t> out;
  auto emit = [&](int M, int K, bool transposed) {
    for (int mt = 0; mt < M / 16; ++mt)
      for (int kt = 0; kt < K / 16; ++kt)
        for (int part = 0; part < 2; ++part)
          for (int l = 0; l < 64; ++l) {
            uint64_t v = 0;
            for (int q = 0; q < 4; ++q) {
              const int row = mt * 16 + l % 16, col = kt * 16 + 4 * (l / 16) + q;
              const double x = transposed ? (row < np ? dk(col, row) : 0.0) : dk(row, col);
              const uint16_t hi = bf16_bits((float)x);
              const uint16_t bits = part == 0 ? hi : bf16_bits((float)(x - (double)bf16_value(hi)));
              v |= (uint64_t)bits << (16 * q);
            }
            out.push_back(v);
          }
  };
  emit(ru(2 * np), ru(np), false);
  emit(ru(np), ru(2 * np), true);
  return out;
}

// [n_nodes x nf] vectors between the caller's node order and the internal (patch) order
void to_internal(const sem_ctx* c, const double* ext, double* in) {
  const int nf = c->nf;
  for (int64_t k = 0; k < c->n_nodes; ++k)
    std::copy(ext + k * nf, ext + (k + 1) * nf, in + (int64_t)c->plan.node_new[k] * nf);
}

void to_external(const sem_ctx* c, const double* in, double* ext) {
  const int nf = c->nf;
  for (int64_t k = 0; k < c->n_nodes; ++k)
    std::copy(in + (int64_t)c->plan.node_new[k] * nf, in + (int64_t)(c->plan.node_new[k] + 1) * nf, ext + k * nf);
}

// Attaches every source to the first occurrence (in assembly order) of its node and
// uploads the element-local lists and the signals [n_sig][nft] into `out`.
int upload_sources(sem_ctx* c, int64_t n_src, const int64_t* dofs, const double* amp, const int32_t* sig,
                   int n_sig, int64_t nft, const double* st, SourceSet& out) {
  std::vector<std::vector<std::pair<int32_t, int64_t>>> per_el(c->n_el);
  for (int64_t s = 0; s < n_src; ++s) {
    if (dofs[s] < 0 || dofs[s] >= c->dof) return fail(c, SEM_E_ARG, "excitation dof out of range");
    if (sig && (sig[s] < 0 || sig[s] >= n_sig)) return fail(c, SEM_E_ARG, "excitation signal index out of range");
    const int64_t node = dofs[s] / c->nf;
    const int f = (int)(dofs[s] % c->nf);
    if (c->rowptr[node] == c->rowptr[node + 1]) return fail(c, SEM_E_ARG, "excited node belongs to no element");
    const int32_t q = c->occ[c->rowptr[node]];
    per_el[c->inv_perm[q / c->nloc]].push_back({(int32_t)((q % c->nloc) * c->nf + f), s});
  }
  std::vector<int32_t> ptr(c->n_el + 1, 0), pf, sg;
  std::vector<double> am;
  for (int64_t e = 0; e < c->n_el; ++e) {
    for (auto& pr : per_el[e]) {
      pf.push_back(pr.first);
      sg.push_back(sig ? sig[pr.second] : 0);
      am.push_back(amp[pr.second]);
    }
    ptr[e + 1] = (int32_t)pf.size();
  }
  if (!pf.empty()) {
    HIPCHK(c, dupload(out.ptr, ptr.data(), ptr.size(), c->stream));
    HIPCHK(c, dupload(out.pf, pf.data(), pf.size(), c->stream));
    HIPCHK(c, dupload(out.sig, sg.data(), sg.size(), c->stream));
    HIPCHK(c, dupload(out.amp, am.data(), am.size(), c->stream));
  }
  HIPCHK(c, dupload(out.st, st, (size_t)n_sig * nft, c->stream));
  out.nft = nft;
  return SEM_OK;
}

}  // namespace

extern "C" {

const char* sem_version(void) { return "semhip 0.1 (gfx950)"; }

const char* sem_last_error(const sem_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int sem_create(sem_ctx** out, int device) {
  if (!out) return SEM_E_ARG;
  *out = nullptr;
  sem_ctx* c = new (std::nothrow) sem_ctx();
  if (!c) return SEM_E_NOMEM;
  c->device = device;
  hipError_t e = hipSetDevice(device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipMalloc((void**)&c->d_ctr, semk::kCtr * sizeof(uint64_t));
  if (e == hipSuccess) e = hipMemset(c->d_ctr, 0, semk::kCtr * sizeof(uint64_t));
  if (e != hipSuccess) {
    if (c->d_ctr) hipFree(c->d_ctr);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
    return SEM_E_HIP;
  }
  // the allocation-free entry point's scratch on this device (a failed lookup leaves that
  // entry point refusing; everything else works)
  hipError_t se;
  c->scratch_dev = scratch_on(device, &se);
  *out = c;
  return SEM_OK;
}

int sem_destroy(sem_ctx* c) {
  if (!c) return SEM_OK;
  hipSetDevice(c->device);
  if (c->stream) hipStreamSynchronize(c->stream);
  drop_graph(c);
  dfree(c->d_gtab); dfree(c->d_geom); dfree(c->d_geomf); dfree(c->d_node_new);
  dfree(c->d_el0); dfree(c->d_pn_base); dfree(c->d_pn_owned); dfree(c->d_pn_id); dfree(c->d_slot_base);
  dfree(c->d_occ_base); dfree(c->d_sh_id); dfree(c->d_sh_ptr); dfree(c->d_sh_slot);
  dfree(c->d_occ_ptr); dfree(c->d_pocc); dfree(c->d_lmap); dfree(c->d_mtab); dfree(c->d_partial);
  dfree(c->d_optr_base); dfree(c->d_lm_base); dfree(c->d_occ8); dfree(c->d_o8_base);
  dfree(c->d_mat); dfree(c->d_elem_mat); dfree(c->d_sscale); dfree(c->d_minv);
  for (auto& p : c->d_u) dfree(p);
  c->ex.release();
  dfree(c->d_out_ptr); dfree(c->d_out_w); dfree(c->d_traces);
  dfree(c->d_tp_ptr); dfree(c->d_tp_lf); dfree(c->d_tp_t); dfree(c->d_ts_ptr); dfree(c->d_ts_f);
  dfree(c->d_ts_t); dfree(c->d_ring);
  dfree(c->d_frame_nn); dfree(c->d_frame_tab); dfree(c->d_frames); dfree(c->d_ctr);
  if (c->stream) hipStreamDestroy(c->stream);
  delete c;
  return SEM_OK;
}

}  // extern "C"

namespace {

int upload_body(sem_ctx* c, int kind, int n_xy, int n_z, int64_t n_el, const int32_t* nodes, int64_t n_nodes,
                const double* coords, const int32_t* group) {
  if (!c) return SEM_E_ARG;
  if (!nodes || !coords || n_el <= 0 || n_nodes <= 0) return fail(c, SEM_E_ARG, "null or empty mesh");
  if (kind == SEM_SHELL5) {
    if (!semk::shell_supported(n_xy)) return fail(c, SEM_E_UNSUPPORTED, "shell: n_xy must be 3..8");
    n_z = 1;
  } else if (kind == SEM_SOLID3) {
    if (!semk::solid_supported(n_xy, n_z))
      return fail(c, SEM_E_UNSUPPORTED, "solid: supported (n_xy,n_z) are n_xy in 3..6 with n_z = 3 or n_z = n_xy");
  } else {
    return fail(c, SEM_E_ARG, "unknown element kind");
  }
  HIPCHK(c, hipSetDevice(c->device));
  drop_graph(c);
  c->kind = kind;
  c->nf = kind == SEM_SHELL5 ? 5 : 3;
  c->nxy = n_xy;
  c->nz = n_z;
  c->nloc = n_xy * n_xy * n_z;
  c->n_el = n_el;
  c->n_nodes = n_nodes;
  c->dof = n_nodes * c->nf;
  const int64_t L = n_el * c->nloc;
  if (L > INT32_MAX || c->dof > ((int64_t)1 << 40)) return fail(c, SEM_E_ARG, "mesh too large for int32 point ids");
  for (int64_t q = 0; q < L; ++q)
    if (nodes[q] < 0 || nodes[q] >= n_nodes) return fail(c, SEM_E_ARG, "node id out of range at point " + std::to_string(q));
  c->h_nodes.assign(nodes, nodes + L);
  c->h_coords.assign(coords, coords + 3 * n_nodes);
  // GLL tables
  std::vector<double> x, w, D, xz, wz, Dz;
  semhost::gll(n_xy, x, w);
  semhost::lagrange_derivative(n_xy, x, D);
  std::vector<double> gtab(D);
  if (kind == SEM_SOLID3) {
    semhost::gll(n_z, xz, wz);
    semhost::lagrange_derivative(n_z, xz, Dz);
    gtab.insert(gtab.end(), Dz.begin(), Dz.end());
    gtab.insert(gtab.end(), w.begin(), w.end());
    gtab.insert(gtab.end(), wz.begin(), wz.end());
  } else {
    gtab.insert(gtab.end(), w.begin(), w.end());
  }
  {  // float copy of the tables after the doubles
    const size_t nd = gtab.size();
    std::vector<float> gf(gtab.begin(), gtab.end());
    gtab.resize(nd + (nd + 1) / 2, 0.0);
    std::memcpy(gtab.data() + nd, gf.data(), gf.size() * sizeof(float));
  }
  if (kind == SEM_SOLID3) {
    // even-odd tables of D, D^T, Dz, Dz^T (semk::eo_table layout), doubles then their float copy
    std::vector<double> eo;
    for (int q = 0; q < 4; ++q) {
      const std::vector<double>& M0 = q < 2 ? D : Dz;
      const int n = q < 2 ? n_xy : n_z;
      std::vector<double> M(M0);
      if (q % 2)
        for (int i = 0; i < n; ++i)
          for (int m = 0; m < n; ++m) M[i * n + m] = M0[m * n + i];
      semhost::even_odd_table(n, M, eo);
    }
    const size_t ne = eo.size();
    std::vector<float> ef(eo.begin(), eo.end());
    eo.resize(ne + (ne + 1) / 2, 0.0);
    std::memcpy(eo.data() + ne, ef.data(), ef.size() * sizeof(float));
    gtab.insert(gtab.end(), eo.begin(), eo.end());
  }
  HIPCHK(c, dupload(c->d_gtab, gtab.data(), gtab.size(), c->stream));
  dfree(c->d_mtab);
  if (kind == SEM_SHELL5 && semk::shell_mfma_supported(n_xy)) {
    std::vector<uint64_t> mt = mfma_tables(n_xy, D);
    HIPCHK(c, dupload(c->d_mtab, mt.data(), mt.size(), c->stream));
  }
  // geometry, affine detection
  const int G = kind == SEM_SHELL5 ? 5 : 10;
  std::vector<double> pg;
  std::vector<double> all((size_t)L * G);
  bool affine = true;
  for (int64_t e = 0; e < n_el; ++e) {
    element_geometry(c, e, D, Dz, pg);
    double scale = 0;
    for (int q = 0; q < G; ++q) scale = std::max(scale, std::fabs(pg[q]));
    for (int p = 1; p < c->nloc && affine; ++p)
      for (int q = 0; q < G; ++q)
        if (std::fabs(pg[(size_t)p * G + q] - pg[q]) > 1e-11 * scale) { affine = false; break; }
    for (int p = 0; p < c->nloc; ++p)
      if (pg[(size_t)p * G + G - 1] <= 0) return fail(c, SEM_E_ARG, "non-positive Jacobian determinant in element " + std::to_string(e));
    std::copy(pg.begin(), pg.end(), all.begin() + (size_t)e * c->nloc * G);
  }
  c->affine = affine;
  c->geom_stride = G;
  // axis-aligned solid elements: the off-diagonal inverse-Jacobian entries are the rounding
  // of the geometry sums (about 1e-13 of the diagonal); below 1e-12 of the element's largest
  // entry the staged kernel drops them
  c->diag = false;
  if (affine && kind == SEM_SOLID3) {
    c->diag = true;
    for (int64_t e = 0; e < n_el && c->diag; ++e) {
      const double* g = &all[(size_t)e * c->nloc * G];
      double big = 0;
      for (int q = 0; q < 9; ++q) big = std::max(big, std::fabs(g[q]));
      for (int q = 0; q < 9; ++q)
        if (q % 4 != 0 && std::fabs(g[q]) > 1e-12 * big) { c->diag = false; break; }
    }
  }
  // per-point quadrature weight x detJ
  c->h_wdetj.resize(L);
  for (int64_t e = 0; e < n_el; ++e)
    for (int p = 0; p < c->nloc; ++p) {
      const int i = p % n_xy, j = (p / n_xy) % n_xy, k = p / (n_xy * n_xy);
      double ww = w[i] * w[j] * (kind == SEM_SOLID3 ? wz[k] : 1.0);
      c->h_wdetj[e * c->nloc + p] = ww * all[((size_t)e * c->nloc + p) * G + G - 1];
    }
  // index maps
  if (!semhost::basket_colouring(n_el, c->nloc, nodes, n_nodes, c->basket, c->brow, c->rows))
    return fail(c, SEM_E_ARG, "a node occurs in more than 12 element points");
  semhost::inverse_map(n_el, c->nloc, nodes, n_nodes, c->basket, c->rowptr, c->occ);
  const int P = c->patch_req > 0 ? c->patch_req
                                  : (kind == SEM_SHELL5 ? semk::shell_patch_elems(n_xy, false)
                                                        : (c->variant == 1 || c->variant == 3
                                                               ? 8
                                                               : semk::solid_patch_elems(n_xy, n_z, false)));
  if (!semk::variant_accepts(kind, c->variant, P))
    return fail(c, SEM_E_UNSUPPORTED, "element-kernel variant " + std::to_string(c->variant) + " does not take " +
                                          std::to_string(P) + "-element patches");
  if (!semhost::build_patches(n_el, c->nloc, nodes, n_nodes, coords, P, c->basket, c->plan, group))
    return fail(c, SEM_E_ARG, "patch construction failed");
  const semhost::PatchPlan& pl = c->plan;
  c->inv_perm.assign(n_el, 0);
  for (int64_t e = 0; e < n_el; ++e) c->inv_perm[pl.perm[e]] = (int32_t)e;
  // geometry in patch order: per element (affine) or per point
  if (affine) {
    c->h_geom.resize((size_t)n_el * G);
    for (int64_t e = 0; e < n_el; ++e) {
      const size_t src = (size_t)pl.perm[e] * c->nloc * G;
      std::copy(all.begin() + src, all.begin() + src + G, c->h_geom.begin() + (size_t)e * G);
    }
  } else {
    c->h_geom.resize((size_t)L * G);
    for (int64_t e = 0; e < n_el; ++e)
      for (int p = 0; p < c->nloc; ++p) {
        const size_t src = ((size_t)pl.perm[e] * c->nloc + p) * G, dst = ((size_t)e * c->nloc + p) * G;
        std::copy(all.begin() + src, all.begin() + src + G, c->h_geom.begin() + dst);
        c->h_geom[dst + G - 1] = c->h_wdetj[(size_t)pl.perm[e] * c->nloc + p];
      }
  }
  // the FP32 form needs the least LDS; each run checks its own form (kernel_ready)
  const size_t lds = semk::patch_lds_bytes(kind, n_xy, n_z, true, pl.max_pn, c->variant, P);
  if (lds > 160 * 1024) return fail(c, SEM_E_UNSUPPORTED, "patch needs more LDS than a CU has");
  HIPCHK(c, dupload(c->d_geom, c->h_geom.data(), c->h_geom.size(), c->stream));
  dfree(c->d_geomf);
  if (!affine) {
    std::vector<float> gf(c->h_geom.begin(), c->h_geom.end());
    HIPCHK(c, dupload(c->d_geomf, gf.data(), gf.size(), c->stream));
  }
  HIPCHK(c, dupload(c->d_node_new, pl.node_new.data(), pl.node_new.size(), c->stream));
  c->n_used = 0;
  for (int32_t nd : pl.pn_id) c->n_used = std::max<int64_t>(c->n_used, (int64_t)nd + 1);
  HIPCHK(c, dupload(c->d_el0, pl.el0.data(), pl.el0.size(), c->stream));
  HIPCHK(c, dupload(c->d_pn_base, pl.pn_base.data(), pl.pn_base.size(), c->stream));
  HIPCHK(c, dupload(c->d_pn_owned, pl.pn_owned.data(), pl.pn_owned.size(), c->stream));
  HIPCHK(c, dupload(c->d_pn_id, pl.pn_id.data(), pl.pn_id.size(), c->stream));
  HIPCHK(c, dupload(c->d_slot_base, pl.slot_base.data(), pl.slot_base.size(), c->stream));
  HIPCHK(c, dupload(c->d_occ_ptr, pl.occ_ptr.data(), pl.occ_ptr.size(), c->stream));
  HIPCHK(c, dupload(c->d_occ_base, pl.occ_base.data(), pl.occ_base.size(), c->stream));
  HIPCHK(c, dupload(c->d_pocc, pl.occ.data(), pl.occ.size(), c->stream));
  HIPCHK(c, dupload(c->d_lmap, pl.lmap.data(), pl.lmap.size(), c->stream));
  HIPCHK(c, dupload(c->d_optr_base, pl.optr_base.data(), pl.optr_base.size(), c->stream));
  HIPCHK(c, dupload(c->d_lm_base, pl.lm_base.data(), pl.lm_base.size(), c->stream));
  HIPCHK(c, dupload(c->d_occ8, pl.occ8.data(), pl.occ8.size(), c->stream));
  HIPCHK(c, dupload(c->d_o8_base, pl.o8_base.data(), pl.o8_base.size(), c->stream));
  dfree(c->d_sh_id); dfree(c->d_sh_ptr); dfree(c->d_sh_slot);
  if (!pl.sh_id.empty()) {
    HIPCHK(c, dupload(c->d_sh_id, pl.sh_id.data(), pl.sh_id.size(), c->stream));
    HIPCHK(c, dupload(c->d_sh_ptr, pl.sh_ptr.data(), pl.sh_ptr.size(), c->stream));
    HIPCHK(c, dupload(c->d_sh_slot, pl.sh_slot.data(), pl.sh_slot.size(), c->stream));
  }
  dfree(c->d_partial);
  HIPCHK(c, hipMalloc((void**)&c->d_partial, (size_t)std::max<int64_t>(pl.n_slots, 1) * c->nf * sizeof(double)));
  // the patch kernels derive owned ids as pn_id[pn0] + ln
  for (int64_t b = 0; b < pl.n_patch; ++b)
    for (int32_t ln = 1; ln < pl.pn_owned[b]; ++ln)
      if (pl.pn_id[pl.pn_base[b] + ln] != pl.pn_id[pl.pn_base[b]] + ln)
        return fail(c, SEM_E_STATE, "patch plan: owned nodes not numbered consecutively");
  // which kernel finalises each node: patch (owned) or boundary list
  c->owner_patch.assign(n_nodes, -1);
  c->owner_ln.assign(n_nodes, -1);
  c->shared_idx.assign(n_nodes, -1);
  for (int64_t b = 0; b < pl.n_patch; ++b)
    for (int32_t ln = 0; ln < pl.pn_owned[b]; ++ln) {
      const int32_t nd = pl.pn_id[pl.pn_base[b] + ln];
      c->owner_patch[nd] = (int32_t)b;
      c->owner_ln[nd] = ln;
    }
  for (size_t k = 0; k < pl.sh_id.size(); ++k) c->shared_idx[pl.sh_id[k]] = (int32_t)k;
  for (auto& p : c->d_u) {
    dfree(p);
    HIPCHK(c, hipMalloc((void**)&p, c->dof * sizeof(double)));
    HIPCHK(c, hipMemsetAsync(p, 0, c->dof * sizeof(double), c->stream));
  }
  c->rot = 0;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->have_material = false;
  c->ex.release();
  c->n_out = 0;
  c->n_terms = 0;
  c->n_frames = 0;
  c->n_ftab = 0;
  dfree(c->d_frame_tab);
  dfree(c->d_tp_ptr); dfree(c->d_ts_ptr);
  return SEM_OK;
}

// A refused upload (bad ids, inverted elements, unsupported sizes, device errors) leaves the
// context without a mesh and material: the checks run while the new mesh is half set up, so
// nothing of the old one or the refused one stays in force, and later calls report a state
// error instead of running on mismatched buffers.
int upload_impl(sem_ctx* c, int kind, int n_xy, int n_z, int64_t n_el, const int32_t* nodes, int64_t n_nodes,
                const double* coords, const int32_t* group) {
  const int rc = upload_body(c, kind, n_xy, n_z, n_el, nodes, n_nodes, coords, group);
  if (rc != SEM_OK && c) {
    c->kind = 0;
    c->have_material = false;
    drop_graph(c);
  }
  return rc;
}

}  // namespace

extern "C" {

int sem_upload_mesh(sem_ctx* c, int kind, int n_xy, int n_z, int64_t n_el, const int32_t* nodes,
                    int64_t n_nodes, const double* coords) {
  return upload_impl(c, kind, n_xy, n_z, n_el, nodes, n_nodes, coords, nullptr);
}

int sem_upload_batch(sem_ctx* c, int kind, int n_xy, int n_z, int n_cases, const int64_t* el_ptr,
                     const int32_t* nodes, const int64_t* node_ptr, const double* coords) {
  if (!c) return SEM_E_ARG;
  if (n_cases <= 0 || !el_ptr || !node_ptr || !nodes || !coords || el_ptr[0] != 0 || node_ptr[0] != 0)
    return fail(c, SEM_E_ARG, "bad batch description");
  const int nloc = kind == SEM_SHELL5 ? n_xy * n_xy : n_xy * n_xy * n_z;
  for (int k = 0; k < n_cases; ++k)
    if (el_ptr[k + 1] < el_ptr[k] || node_ptr[k + 1] < node_ptr[k]) return fail(c, SEM_E_ARG, "batch offsets must ascend");
  const int64_t n_el = el_ptr[n_cases], n_nodes = node_ptr[n_cases];
  if (n_el * nloc > INT32_MAX || n_nodes > INT32_MAX) return fail(c, SEM_E_ARG, "batch too large for int32 ids");
  std::vector<int32_t> un((size_t)n_el * nloc), grp((size_t)n_el);
  for (int k = 0; k < n_cases; ++k) {
    const int64_t nn = node_ptr[k + 1] - node_ptr[k];
    for (int64_t e = el_ptr[k]; e < el_ptr[k + 1]; ++e) {
      grp[e] = k;
      for (int p = 0; p < nloc; ++p) {
        const int32_t v = nodes[e * nloc + p];
        if (v < 0 || v >= nn) return fail(c, SEM_E_ARG, "node id out of range in batch case " + std::to_string(k));
        un[(size_t)e * nloc + p] = (int32_t)(node_ptr[k] + v);
      }
    }
  }
  return upload_impl(c, kind, n_xy, n_z, n_el, un.data(), n_nodes, coords, grp.data());
}

int sem_set_kernel(sem_ctx* c, int variant, int patch_elems) {
  if (!c || variant < 0 || variant > 3 || patch_elems < 0) return SEM_E_ARG;
  drop_graph(c);
  c->variant = variant;
  c->patch_req = patch_elems;
  return SEM_OK;
}

int sem_get_kernel(const sem_ctx* c, int* variant, int* patch_elems) {
  if (!c || !variant || !patch_elems) return SEM_E_ARG;
  *variant = c->variant;
  *patch_elems = c->kind ? c->plan.P : c->patch_req;
  return SEM_OK;
}

int sem_basket_rows(const sem_ctx* c, int64_t* rows) {
  if (!c || !rows) return SEM_E_ARG;
  if (!c->kind) return SEM_E_STATE;
  *rows = c->rows;
  return SEM_OK;
}

int sem_get_baskets(const sem_ctx* c, int32_t* I_G, int32_t* I_L) {
  if (!c || !I_G || !I_L) return SEM_E_ARG;
  if (!c->kind) return SEM_E_STATE;
  const int64_t L = c->n_el * c->nloc;
  std::fill(I_G, I_G + 12 * c->rows, 0);
  std::fill(I_L, I_L + 12 * c->rows, 0);
  for (int64_t q = 0; q < L; ++q) {
    const int64_t at = (int64_t)c->basket[q] * c->rows + c->brow[q];
    I_G[at] = c->h_nodes[q] + 1;
    I_L[at] = (int32_t)(q + 1);
  }
  return SEM_OK;
}

int sem_mesh_is_affine(const sem_ctx* c, int* affine) {
  if (!c || !affine) return SEM_E_ARG;
  if (!c->kind) return SEM_E_STATE;
  *affine = c->affine ? 1 : 0;
  return SEM_OK;
}

int sem_set_material(sem_ctx* c, int n_mat, int n_coef, const double* coef, const int32_t* elem_mat,
                     const double* stiff_scale, const double* rho_scale) {
  if (!c) return SEM_E_ARG;
  if (!c->kind) return fail(c, SEM_E_STATE, "upload a mesh first");
  const int need = c->kind == SEM_SHELL5 ? 23 : 14;
  if (n_mat <= 0 || n_coef != need || !coef || !elem_mat)
    return fail(c, SEM_E_ARG, "material table must be n_mat x " + std::to_string(need));
  for (int64_t e = 0; e < c->n_el; ++e)
    if (elem_mat[e] < 0 || elem_mat[e] >= n_mat) return fail(c, SEM_E_ARG, "elem_mat out of range");
  HIPCHK(c, hipSetDevice(c->device));
  drop_graph(c);
  c->n_coef = n_coef;
  HIPCHK(c, dupload(c->d_mat, coef, (size_t)n_mat * n_coef, c->stream));
  {
    std::vector<int32_t> em(c->n_el);
    for (int64_t e = 0; e < c->n_el; ++e) em[e] = elem_mat[c->plan.perm[e]];
    HIPCHK(c, dupload(c->d_elem_mat, em.data(), em.size(), c->stream));
  }
  dfree(c->d_sscale);
  if (stiff_scale) {
    std::vector<double> ss(c->n_el);
    for (int64_t e = 0; e < c->n_el; ++e) ss[e] = stiff_scale[c->plan.perm[e]];
    HIPCHK(c, dupload(c->d_sscale, ss.data(), ss.size(), c->stream));
  }
  // lumped mass, assembled per node in basket order (force_excit_c.m:497-503;
  // mass3D_v2.m:51-57 + solid engine :129-145)
  const int nf = c->nf, np = c->nloc;
  std::vector<double> mloc((size_t)np * nf);
  c->h_mass.assign(c->dof, 0.0);
  std::vector<double> mq((size_t)c->n_el * np * nf);
  for (int64_t e = 0; e < c->n_el; ++e) {
    const double* m = coef + (int64_t)elem_mat[e] * n_coef;
    const double rs = rho_scale ? rho_scale[e] : 1.0;
    for (int p = 0; p < np; ++p) {
      const double W = c->h_wdetj[e * np + p];
      double* o = &mq[((size_t)e * np + p) * nf];
      if (c->kind == SEM_SHELL5) {
        const double m11 = m[21] * rs, I11 = m[22] * rs;
        o[0] = W * m11; o[1] = W * I11; o[2] = W * m11; o[3] = W * I11; o[4] = W * m11;
      } else {
        const double rho = m[13] * rs;
        o[0] = W * rho; o[1] = W * rho; o[2] = W * rho;
      }
    }
  }
  // fields that share a mass coefficient get bitwise equal sums (same terms, same order), so
  // the inverse mass is kept once per node and coefficient
  const int nm = semk::mass_cols(nf);
  std::vector<double> minv((size_t)c->n_nodes * nm, 0.0);
  for (int64_t k = 0; k < c->n_nodes; ++k)
    for (int f = 0; f < nf; ++f) {
      double s = 0.0;
      for (int32_t o = c->rowptr[k]; o < c->rowptr[k + 1]; ++o) s += mq[(size_t)c->occ[o] * nf + f];
      c->h_mass[k * nf + f] = s;
      if (f >= nm && s != c->h_mass[k * nf + semk::mass_col(nf, f)])
        return fail(c, SEM_E_STATE, "fields of one node with different lumped masses");
      if (f < nm) minv[(int64_t)c->plan.node_new[k] * nm + semk::mass_col(nf, f)] = s != 0.0 ? 1.0 / s : 0.0;
    }
  HIPCHK(c, dupload(c->d_minv, minv.data(), minv.size(), c->stream));
  c->have_material = true;
  return SEM_OK;
}

int sem_set_damping(sem_ctx* c, const double* dv_field) {
  if (!c || !dv_field) return SEM_E_ARG;
  if (!c->kind) return fail(c, SEM_E_STATE, "upload a mesh first");
  drop_graph(c);
  for (int f = 0; f < c->nf; ++f) c->dv[f] = dv_field[f];
  return SEM_OK;
}

int sem_get_mass(const sem_ctx* c, double* mass) {
  if (!c || !mass) return SEM_E_ARG;
  if (!c->have_material) return SEM_E_STATE;
  std::copy(c->h_mass.begin(), c->h_mass.end(), mass);
  return SEM_OK;
}

int sem_set_excitation_multi(sem_ctx* c, int64_t n_src, const int64_t* dofs, const double* amp,
                             const int32_t* sig, int n_sig, int64_t nft, const double* st) {
  if (!c) return SEM_E_ARG;
  if (!c->kind) return fail(c, SEM_E_STATE, "upload a mesh first");
  if (n_src < 0 || (n_src > 0 && (!dofs || !amp)) || n_sig < 1 || nft <= 0 || !st)
    return fail(c, SEM_E_ARG, "bad excitation");
  // the same checks as the upload, before the old sources go: a refused call leaves them
  for (int64_t k = 0; k < n_src; ++k) {
    if (dofs[k] < 0 || dofs[k] >= c->dof) return fail(c, SEM_E_ARG, "excitation dof out of range");
    if (sig && (sig[k] < 0 || sig[k] >= n_sig)) return fail(c, SEM_E_ARG, "excitation signal index out of range");
    const int64_t node = dofs[k] / c->nf;
    if (c->rowptr[node] == c->rowptr[node + 1]) return fail(c, SEM_E_ARG, "excited node belongs to no element");
  }
  HIPCHK(c, hipSetDevice(c->device));
  drop_graph(c);
  c->ex.release();
  return upload_sources(c, n_src, dofs, amp, sig, n_sig, nft, st, c->ex);
}

int sem_set_excitation(sem_ctx* c, int64_t n_src, const int64_t* dofs, const double* amp, int64_t nft,
                       const double* st) {
  return sem_set_excitation_multi(c, n_src, dofs, amp, nullptr, 1, nft, st);
}

int sem_shell_pzt_sensor(const sem_ctx* c, int64_t n_sel, const int32_t* elems, double t_pzt, double h, double e31,
                         double e32, double g33, double* w) {
  if (!c || !elems || !w || n_sel <= 0 || g33 == 0.0) return SEM_E_ARG;
  if (c->kind != SEM_SHELL5) return SEM_E_STATE;
  const int n = c->nxy, np = c->nloc;
  std::vector<double> x, wq, D, Dz, g;
  semhost::gll(n, x, wq);
  semhost::lagrange_derivative(n, x, D);
  std::fill(w, w + c->dof, 0.0);
  // voltage = mean over the sensor's points of t/g33 (e31 (u_x + h/2 phix_x) + e32 (v_y + h/2 phiy_y)),
  // u_x = (d/dxi u) invJ11 + (d/deta u) invJ12, v_y = (d/dxi v) invJ21 + (d/deta v) invJ22
  const double cf = t_pzt / g33 / (double)(n_sel * np);
  for (int64_t s = 0; s < n_sel; ++s) {
    const int64_t e = elems[s];
    if (e < 0 || e >= c->n_el) return SEM_E_ARG;
    element_geometry(c, e, D, Dz, g);
    const int32_t* nd = c->h_nodes.data() + e * np;
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i) {
        const double* gp = &g[(size_t)(j * n + i) * 5];
        for (int k = 0; k < n; ++k) {
          const int64_t a = nd[j * n + k], b = nd[k * n + i];  // xi line, eta line
          const double cu_a = cf * e31 * D[i * n + k] * gp[0], cu_b = cf * e31 * D[j * n + k] * gp[1];
          const double cv_a = cf * e32 * D[i * n + k] * gp[2], cv_b = cf * e32 * D[j * n + k] * gp[3];
          w[a * 5 + 0] += cu_a;
          w[b * 5 + 0] += cu_b;
          w[a * 5 + 1] += h / 2 * cu_a;
          w[b * 5 + 1] += h / 2 * cu_b;
          w[a * 5 + 2] += cv_a;
          w[b * 5 + 2] += cv_b;
          w[a * 5 + 3] += h / 2 * cv_a;
          w[b * 5 + 3] += h / 2 * cv_b;
        }
      }
  }
  return SEM_OK;
}

int sem_set_outputs(sem_ctx* c, int n_out, const int64_t* ptr, const int64_t* dofs, const double* w,
                    int quantity) {
  if (!c) return SEM_E_ARG;
  if (!c->kind) return fail(c, SEM_E_STATE, "upload a mesh first");
  if (n_out < 0 || quantity < 0 || quantity > 3) return fail(c, SEM_E_ARG, "bad outputs");
  // every check runs before the context changes: a refused call leaves the old outputs
  int64_t nnz = 0;
  if (n_out > 0) {
    if (!ptr || !dofs || !w) return fail(c, SEM_E_ARG, "null output arrays");
    if (ptr[0] != 0) return fail(c, SEM_E_ARG, "ptr[0] must be 0");
    for (int j = 0; j < n_out; ++j)
      if (ptr[j + 1] < ptr[j]) return fail(c, SEM_E_ARG, "ptr must be non-decreasing");
    nnz = ptr[n_out];
    if (nnz > INT32_MAX) return fail(c, SEM_E_ARG, "too many output terms");
    for (int64_t q = 0; q < nnz; ++q)
      if (dofs[q] < 0 || dofs[q] >= c->dof) return fail(c, SEM_E_ARG, "output dof out of range");
  }
  // route every term to the kernel that finalises its node
  const semhost::PatchPlan& pl = c->plan;
  std::vector<std::vector<std::pair<int32_t, int32_t>>> per_patch(pl.n_patch), per_shared(pl.sh_id.size());
  for (int64_t t = 0; t < nnz; ++t) {
    const int64_t nd = c->plan.node_new[dofs[t] / c->nf];
    const int32_t f = (int32_t)(dofs[t] % c->nf);
    if (c->owner_patch[nd] >= 0)
      per_patch[c->owner_patch[nd]].push_back({c->owner_ln[nd] * c->nf + f, (int32_t)t});
    else if (c->shared_idx[nd] >= 0)
      per_shared[c->shared_idx[nd]].push_back({f, (int32_t)t});
    else
      return fail(c, SEM_E_ARG, "output dof on a node that belongs to no element");
  }
  HIPCHK(c, hipSetDevice(c->device));
  drop_graph(c);
  c->n_out = 0;
  c->n_terms = 0;
  c->out_q = quantity;
  dfree(c->d_out_ptr); dfree(c->d_out_w);
  dfree(c->d_tp_ptr); dfree(c->d_tp_lf); dfree(c->d_tp_t); dfree(c->d_ts_ptr); dfree(c->d_ts_f);
  dfree(c->d_ts_t); dfree(c->d_ring);
  if (n_out == 0) return SEM_OK;
  std::vector<int32_t> tp_ptr(pl.n_patch + 1, 0), tp_lf, tp_t, ts_ptr(pl.sh_id.size() + 1, 0), ts_f, ts_t;
  for (int64_t b = 0; b < pl.n_patch; ++b) {
    for (auto& x : per_patch[b]) { tp_lf.push_back(x.first); tp_t.push_back(x.second); }
    tp_ptr[b + 1] = (int32_t)tp_lf.size();
  }
  for (size_t k = 0; k < per_shared.size(); ++k) {
    for (auto& x : per_shared[k]) { ts_f.push_back(x.first); ts_t.push_back(x.second); }
    ts_ptr[k + 1] = (int32_t)ts_f.size();
  }
  c->h_out_ptr.assign(ptr, ptr + n_out + 1);
  c->h_out_dof.assign(dofs, dofs + nnz);
  c->h_out_w.assign(w, w + nnz);
  HIPCHK(c, dupload(c->d_out_ptr, ptr, (size_t)n_out + 1, c->stream));
  if (nnz) HIPCHK(c, dupload(c->d_out_w, w, (size_t)nnz, c->stream));
  if (!tp_lf.empty()) {
    HIPCHK(c, dupload(c->d_tp_ptr, tp_ptr.data(), tp_ptr.size(), c->stream));
    HIPCHK(c, dupload(c->d_tp_lf, tp_lf.data(), tp_lf.size(), c->stream));
    HIPCHK(c, dupload(c->d_tp_t, tp_t.data(), tp_t.size(), c->stream));
  }
  if (!ts_f.empty()) {
    HIPCHK(c, dupload(c->d_ts_ptr, ts_ptr.data(), ts_ptr.size(), c->stream));
    HIPCHK(c, dupload(c->d_ts_f, ts_f.data(), ts_f.size(), c->stream));
    HIPCHK(c, dupload(c->d_ts_t, ts_t.data(), ts_t.size(), c->stream));
  }
  HIPCHK(c, hipMalloc((void**)&c->d_ring, (size_t)kChunk * std::max<int64_t>(nnz, 1) * sizeof(double)));
  c->n_out = n_out;
  c->n_terms = nnz;
  return SEM_OK;
}

int sem_set_frames(sem_ctx* c, int n_frames, const int64_t* frame_nn, int quantity) {
  if (!c) return SEM_E_ARG;
  if (!c->kind) return fail(c, SEM_E_STATE, "upload a mesh first");
  if (n_frames < 0 || (n_frames > 0 && !frame_nn) || quantity < 0 || quantity > 3)
    return fail(c, SEM_E_ARG, "bad frames");
  for (int f = 1; f < n_frames; ++f)
    if (frame_nn[f] <= frame_nn[f - 1]) return fail(c, SEM_E_ARG, "frame steps must ascend");
  if (n_frames && frame_nn[0] < 1) return fail(c, SEM_E_ARG, "frame steps are 1-based");
  // every check above runs before the context changes: a refused call leaves the old frames
  HIPCHK(c, hipSetDevice(c->device));
  drop_graph(c);
  c->n_frames = n_frames;
  c->frame_q = quantity;
  c->h_frame_nn.assign(frame_nn, frame_nn + n_frames);
  dfree(c->d_frame_nn);
  dfree(c->d_frame_tab);
  dfree(c->d_frames);
  c->n_ftab = 0;
  if (n_frames) {
    // step -> frame index table read by the kernels
    c->n_ftab = frame_nn[n_frames - 1];
    std::vector<int32_t> tab(c->n_ftab, -1);
    for (int f = 0; f < n_frames; ++f) tab[frame_nn[f] - 1] = f;
    HIPCHK(c, dupload(c->d_frame_tab, tab.data(), tab.size(), c->stream));
    HIPCHK(c, dupload(c->d_frame_nn, frame_nn, (size_t)n_frames, c->stream));
    HIPCHK(c, hipMalloc((void**)&c->d_frames, (size_t)n_frames * c->dof * sizeof(double)));
  }
  return SEM_OK;
}

int sem_set_guard(sem_ctx* c, double limit, int field, int post) {
  if (!c) return SEM_E_ARG;
  drop_graph(c);
  c->guard_limit = limit;
  c->guard_field = field;
  c->guard_post = post;
  return SEM_OK;
}

int sem_set_state(sem_ctx* c, const double* u, const double* uold) {
  if (!c) return SEM_E_ARG;
  if (!c->kind) return fail(c, SEM_E_STATE, "upload a mesh first");
  HIPCHK(c, hipSetDevice(c->device));
  double* cur = c->d_u[(c->rot + 1) % 3];
  double* old = c->d_u[c->rot];
  const size_t b = c->dof * sizeof(double);
  std::vector<double> h0, h1;
  if (u) {
    h0.resize(c->dof);
    to_internal(c, u, h0.data());
    HIPCHK(c, hipMemcpyAsync(cur, h0.data(), b, hipMemcpyHostToDevice, c->stream));
  } else {
    HIPCHK(c, hipMemsetAsync(cur, 0, b, c->stream));
  }
  if (uold) {
    h1.resize(c->dof);
    to_internal(c, uold, h1.data());
    HIPCHK(c, hipMemcpyAsync(old, h1.data(), b, hipMemcpyHostToDevice, c->stream));
  } else {
    HIPCHK(c, hipMemsetAsync(old, 0, b, c->stream));
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return SEM_OK;
}

int sem_get_state(const sem_ctx* cc, double* u, double* uold) {
  sem_ctx* c = const_cast<sem_ctx*>(cc);
  if (!c) return SEM_E_ARG;
  if (!c->kind) return fail(c, SEM_E_STATE, "upload a mesh first");
  HIPCHK(c, hipSetDevice(c->device));
  const size_t b = c->dof * sizeof(double);
  std::vector<double> h(c->dof);
  if (u) {
    HIPCHK(c, hipMemcpy(h.data(), c->d_u[(c->rot + 1) % 3], b, hipMemcpyDeviceToHost));
    to_external(c, h.data(), u);
  }
  if (uold) {
    HIPCHK(c, hipMemcpy(h.data(), c->d_u[c->rot], b, hipMemcpyDeviceToHost));
    to_external(c, h.data(), uold);
  }
  return SEM_OK;
}

}  // extern "C"

namespace {

// src: the point sources to apply (nullable = none)
semk::PatchArgs patch_args(sem_ctx* c, const SourceSet* src) {
  semk::PatchArgs a{};
  a.el0 = c->d_el0;
  a.pn_base = c->d_pn_base;
  a.pn_owned = c->d_pn_owned;
  a.pn_id = c->d_pn_id;
  a.slot_base = c->d_slot_base;
  a.occ_ptr = c->d_occ_ptr;
  a.optr_base = c->d_optr_base;
  a.occ_base = c->d_occ_base;
  a.occ = c->d_pocc;
  a.occ8 = c->d_occ8;
  a.o8_base = c->d_o8_base;
  a.lm_base = c->d_lm_base;
  a.lmap = c->d_lmap;
  a.max_pn = c->plan.max_pn;
  a.geom = c->d_geom;
  a.geomf = c->d_geomf;
  a.elem_mat = c->d_elem_mat;
  a.mat = c->d_mat;
  a.n_coef = c->n_coef;
  a.sscale = c->d_sscale;
  a.diag = c->diag ? 1 : 0;
  if (src && src->ptr) {
    a.ex_ptr = src->ptr;
    a.ex_pf = src->pf;
    a.ex_sig = src->sig;
    a.ex_amp = src->amp;
    a.st = src->st;
    a.ex_nft = src->nft;
    a.excite = 1;
  }
  a.tp_ptr = c->d_tp_ptr;
  a.tp_lf = c->d_tp_lf;
  a.tp_t = c->d_tp_t;
  a.mtab = c->d_mtab;
  return a;
}

semk::UpdateArgs update_args(sem_ctx* c, const double* Uold, const double* U, double* Unew, double dt,
                             int step_off) {
  semk::UpdateArgs u{};
  u.U = U;
  u.Uold = Uold;
  u.Unew = Unew;
  u.minv = c->d_minv;
  const double a0 = dt > 0 ? 1.0 / (dt * dt) : 0.0, a1 = dt > 0 ? 1.0 / (2.0 * dt) : 0.0;
  for (int f = 0; f < 5; ++f) {
    const double c1 = dt > 0 ? 1.0 / (a0 + a1 * c->dv[f]) : 0.0;
    u.c1[f] = c1;
    u.c2[f] = a0 * c1;
    u.c3[f] = a1 * c->dv[f] * c1;
  }
  u.a0 = a0;
  u.a1 = a1;
  u.step_off = step_off;
  u.frame_tab = c->d_frame_tab;
  u.n_ftab = c->n_ftab;
  u.frame_q = c->frame_q;
  u.frames = c->d_frames;
  u.ctr = c->d_ctr;
  u.guard_limit = c->guard_limit;
  u.guard_field = c->guard_field;
  u.guard_post = c->guard_post;
  u.nan_dof = c->n_nodes > 0 ? (int64_t)c->plan.node_new[0] * c->nf + 2 : -1;
  u.dof = c->dof;
  u.partial = c->d_partial;
  u.out_q = c->out_q;
  u.ring = c->d_ring;
  u.n_terms = c->n_terms;
  return u;
}

semk::SharedArgs shared_args(sem_ctx* c) {
  semk::SharedArgs sa{};
  sa.n_shared = (int64_t)c->plan.sh_id.size();
  sa.sh_id = c->d_sh_id;
  sa.sh_ptr = c->d_sh_ptr;
  sa.sh_slot = c->d_sh_slot;
  sa.ts_ptr = c->d_ts_ptr;
  sa.ts_f = c->d_ts_f;
  sa.ts_t = c->d_ts_t;
  return sa;
}

// one time step at offset `off` of the current chunk, rotation r:
// U^{n-1} = u[r], U^n = u[r+1], U^{n+1} -> u[r+2].  ev (nullable): 3 events.
hipError_t launch_step(sem_ctx* c, int r, int off, double dt, bool fp32, hipEvent_t* ev = nullptr) {
  double* Uold = c->d_u[r % 3];
  double* U = c->d_u[(r + 1) % 3];
  double* Unew = c->d_u[(r + 2) % 3];
  const semk::UpdateArgs ua = update_args(c, Uold, U, Unew, dt, off);
  hipError_t e;
  if (ev) { e = hipEventRecord(ev[0], c->stream); if (e != hipSuccess) return e; }
  e = semk::launch_patch(c->kind, c->nxy, c->nz, fp32, c->affine, 0, c->variant, c->plan.P, c->plan.n_patch, c->d_gtab,
                         patch_args(c, &c->ex), ua, c->stream);
  if (e != hipSuccess) return e;
  if (ev) { e = hipEventRecord(ev[1], c->stream); if (e != hipSuccess) return e; }
  e = semk::launch_shared(c->nf, false, fp32, shared_args(c), ua, c->stream);
  if (e != hipSuccess) return e;
  if (ev) e = hipEventRecord(ev[2], c->stream);
  return e;
}

// One step nn on explicit buffers, without sensor terms, frames or the guard (the
// adjoint's forward and reverse sweeps).  The kernels read the step as ctr[0] + step_off,
// so the caller keeps ctr[0] = 0 and the step goes in step_off.
hipError_t launch_plain_step(sem_ctx* c, const double* Uold, const double* U, double* Unew, int64_t nn, double dt,
                             bool fp32, const SourceSet* src) {
  semk::UpdateArgs ua = update_args(c, Uold, U, Unew, dt, (int)nn);
  ua.frame_tab = nullptr;
  ua.n_ftab = 0;
  ua.frames = nullptr;
  ua.guard_limit = 0.0;
  ua.ring = nullptr;
  ua.n_terms = 0;
  semk::PatchArgs pa = patch_args(c, src);
  pa.tp_ptr = nullptr;
  semk::SharedArgs sa = shared_args(c);
  sa.ts_ptr = nullptr;
  hipError_t e = semk::launch_patch(c->kind, c->nxy, c->nz, fp32, c->affine, 0, c->variant, c->plan.P, c->plan.n_patch, c->d_gtab, pa, ua,
                                    c->stream);
  if (e != hipSuccess) return e;
  return semk::launch_shared(c->nf, false, fp32, sa, ua, c->stream);
}

// Gradient terms of step n: lam = lambda_n, states u_{n-1}, u_n, u_{n+1}.
hipError_t launch_gradient(sem_ctx* c, const double* lam, const double* uo, const double* u, const double* un,
                           double dt, bool fp32, double* gs, double* gr) {
  semk::UpdateArgs ua = update_args(c, uo, u, const_cast<double*>(un), dt, 0);
  ua.lam = lam;
  ua.gs = gs;
  ua.gr = gr;
  ua.gscale_s = -dt * dt;
  ua.gscale_m = -1.0;
  for (int f = 0; f < 5; ++f) ua.gdv[f] = 0.5 * dt * c->dv[f];
  return semk::launch_patch(c->kind, c->nxy, c->nz, fp32, c->affine, 2, c->variant, c->plan.P, c->plan.n_patch, c->d_gtab,
                            patch_args(c, nullptr), ua, c->stream);
}

hipError_t launch_advance(sem_ctx* c, int count, int64_t nn_base, int64_t ld) {
  semk::AdvanceArgs a{};
  a.count = count;
  a.n_out = c->n_out;
  a.out_ptr = c->d_out_ptr;
  a.out_w = c->d_out_w;
  a.ring = c->d_ring;
  a.n_terms = c->n_terms;
  a.traces = c->d_traces;
  a.ld = ld;
  a.nn_base = nn_base;
  a.ctr = c->d_ctr;
  return semk::launch_advance(a, c->stream);
}

// chunk of `count` (<= kChunk) steps followed by the advance kernel; rotation starts at c->rot
hipError_t launch_chunk(sem_ctx* c, int count, double dt, bool fp32, int64_t nn_base, int64_t ld) {
  for (int s = 0; s < count; ++s) {
    hipError_t e = launch_step(c, c->rot + s, s, dt, fp32);
    if (e != hipSuccess) return e;
  }
  return launch_advance(c, count, nn_base, ld);
}

// the chosen element kernel takes the uploaded mesh's patches and fits in LDS
int kernel_ready(sem_ctx* c, bool fp32) {
  if (!semk::variant_accepts(c->kind, c->variant, c->plan.P))
    return fail(c, SEM_E_STATE, "element-kernel variant " + std::to_string(c->variant) + " does not take the mesh's " +
                                    std::to_string(c->plan.P) + "-element patches: upload the mesh again");
  if (c->variant == 2 && (!fp32 || !c->d_mtab))
    return fail(c, SEM_E_UNSUPPORTED, "the matrix-core shell kernel runs the FP32 form of order-4 and order-5 shells");
  if (semk::patch_lds_bytes(c->kind, c->nxy, c->nz, fp32, c->plan.max_pn, c->variant, c->plan.P) > 160 * 1024)
    return fail(c, SEM_E_UNSUPPORTED, std::string("patch needs more LDS than a CU has in the ") +
                                          (fp32 ? "FP32" : "FP64") + " form: use smaller patches");
  return SEM_OK;
}

int prepare_run(sem_ctx* c, double dt, int64_t nn_first, int64_t nsteps, int precision) {
  if (!c) return SEM_E_ARG;
  if (!c->kind || !c->have_material) return fail(c, SEM_E_STATE, "mesh and material must be set");
  if (!(dt > 0) || nn_first < 1 || nsteps < 1) return fail(c, SEM_E_ARG, "bad step range or dt");
  if (precision != SEM_FP64 && precision != SEM_FP32_SPLIT) return fail(c, SEM_E_UNSUPPORTED, "unknown precision");
  if (c->ex.ptr && nn_first + nsteps - 1 > c->ex.nft) return fail(c, SEM_E_ARG, "signal shorter than the last step");
  if (const int rc = kernel_ready(c, precision == SEM_FP32_SPLIT); rc != SEM_OK) return rc;
  HIPCHK(c, hipSetDevice(c->device));
  if (c->n_out > 0 && c->traces_cap < nsteps * c->n_out) {
    dfree(c->d_traces);
    HIPCHK(c, hipMalloc((void**)&c->d_traces, (size_t)nsteps * c->n_out * sizeof(double)));
    c->traces_cap = nsteps * c->n_out;
    drop_graph(c);
  }
  if (c->n_out > 0) HIPCHK(c, hipMemsetAsync(c->d_traces, 0, (size_t)nsteps * c->n_out * sizeof(double), c->stream));
  uint64_t ctr[semk::kCtr] = {(uint64_t)nn_first, UINT64_MAX, 0, 0};
  HIPCHK(c, hipMemcpyAsync(c->d_ctr, ctr, sizeof(ctr), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return SEM_OK;
}

}  // namespace

extern "C" {

int sem_run(sem_ctx* c, double dt, int64_t nn_first, int64_t nn_last, int precision, double* traces,
            int64_t ld_traces, double* frames, double* t_frames, int64_t* nn_done, double* seconds) {
  if (!c) return SEM_E_ARG;
  if (nn_last < nn_first) return fail(c, SEM_E_ARG, "bad step range");
  if (traces && ld_traces < nn_last) return fail(c, SEM_E_ARG, "ld_traces < nn_last");
  const int64_t nsteps = nn_last - nn_first + 1;
  int rc = prepare_run(c, dt, nn_first, nsteps, precision);
  if (rc != SEM_OK) return rc;
  const bool fp32 = precision == SEM_FP32_SPLIT;
  const int64_t ld = nsteps;
  Events ev;
  HIPCHK(c, ev.create(2));
  HIPCHK(c, hipEventRecord(ev[0], c->stream));
  int64_t done = 0;
  const int G = kChunk;  // multiple of 3: the buffer rotation is periodic over a chunk
  if (nsteps >= G) {
    const bool reuse = c->graph && c->graph_rot == c->rot && c->graph_prec == precision && c->graph_dt == dt &&
                       c->graph_base == nn_first && c->graph_ld == ld && c->graph_steps == G;
    if (!reuse) {
      drop_graph(c);
      hipGraph_t g;
      HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
      hipError_t le = launch_chunk(c, G, dt, fp32, nn_first, ld);
      hipError_t ce = hipStreamEndCapture(c->stream, &g);
      if (ce == hipSuccess && le != hipSuccess) hipGraphDestroy(g);
      if (le != hipSuccess) return fail(c, SEM_E_HIP, std::string("capture: ") + hipGetErrorString(le));
      HIPCHK(c, ce);
      hipError_t ie = hipGraphInstantiate(&c->graph, g, nullptr, nullptr, 0);
      hipGraphDestroy(g);
      HIPCHK(c, ie);
      c->graph_rot = c->rot;
      c->graph_prec = precision;
      c->graph_dt = dt;
      c->graph_base = nn_first;
      c->graph_ld = ld;
      c->graph_steps = G;
    }
    // At most about three windows of kGraphWindow graph launches are queued ahead of the
    // host: before marking window k it waits for the end of window k - 2, so a run of tens of
    // thousands of steps does not hand the runtime a thousand launches at once, while the
    // GPU always has at least one window (tens of milliseconds of steps) queued.
    constexpr int64_t kGraphWindow = 32;
    Events win;
    HIPCHK(c, win.create(2));
    int64_t launched = 0;
    while (nsteps - done >= G) {
      HIPCHK(c, hipGraphLaunch(c->graph, c->stream));
      done += G;
      if (++launched % kGraphWindow == 0) {
        const int64_t k = launched / kGraphWindow - 1;
        if (k >= 2) HIPCHK(c, hipEventSynchronize(win[k % 2]));
        HIPCHK(c, hipEventRecord(win[k % 2], c->stream));
      }
    }
  }
  if (done < nsteps) {
    hipError_t e = launch_chunk(c, (int)(nsteps - done), dt, fp32, nn_first, ld);
    if (e != hipSuccess) return fail(c, SEM_E_HIP, std::string("launch: ") + hipGetErrorString(e));
  }
  HIPCHK(c, hipEventRecord(ev[1], c->stream));
  HIPCHK(c, hipEventSynchronize(ev[1]));
  float ms = 0;
  HIPCHK(c, hipEventElapsedTime(&ms, ev[0], ev[1]));
  c->rot = (int)((c->rot + nsteps) % 3);
  if (seconds) *seconds = ms * 1e-3;
  uint64_t ctr[semk::kCtr];
  HIPCHK(c, hipMemcpy(ctr, c->d_ctr, sizeof(ctr), hipMemcpyDeviceToHost));
  const int64_t div_nn = ctr[1] == UINT64_MAX ? 0 : (int64_t)ctr[1];
  const int64_t last = div_nn ? div_nn : nn_last;
  if (traces && c->n_out > 0) {
    std::vector<double> tb((size_t)nsteps * c->n_out);
    HIPCHK(c, hipMemcpy(tb.data(), c->d_traces, tb.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (int j = 0; j < c->n_out; ++j)
      for (int64_t nn = nn_first; nn <= nn_last; ++nn)
        traces[j * ld_traces + nn - 1] = nn <= last ? tb[(size_t)j * nsteps + nn - nn_first] : 0.0;
  }
  // frames of steps in the run; the frame at the tripping step is not kept
  // (force_excit_c.m:712-730 returns before saving it)
  int64_t fi0 = 0;
  while (fi0 < c->n_frames && c->h_frame_nn[fi0] < nn_first) ++fi0;
  int64_t fi_end = fi0;
  while (fi_end < c->n_frames && c->h_frame_nn[fi_end] <= nn_last &&
         (!div_nn || c->h_frame_nn[fi_end] < div_nn))
    ++fi_end;
  if (frames && fi_end > fi0) {
    std::vector<double> h(c->dof);
    for (int64_t f = fi0; f < fi_end; ++f) {
      HIPCHK(c, hipMemcpy(h.data(), c->d_frames + f * c->dof, c->dof * sizeof(double), hipMemcpyDeviceToHost));
      to_external(c, h.data(), frames + f * c->dof);
    }
  }
  if (t_frames)
    for (int64_t f = fi0; f < fi_end; ++f) t_frames[f] = (double)(c->h_frame_nn[f] - 1) * dt;
  if (nn_done) *nn_done = last;
  if (div_nn) return fail(c, SEM_E_DIVERGED, "integration error: divergence guard tripped at step " + std::to_string(div_nn));
  return SEM_OK;
}

int sem_profile(sem_ctx* c, double dt, int64_t nn_first, int64_t nsteps, int precision, double* ms_out) {
  if (!c || !ms_out) return SEM_E_ARG;
  int rc = prepare_run(c, dt, nn_first, nsteps, precision);
  if (rc != SEM_OK) return rc;
  const bool fp32 = precision == SEM_FP32_SPLIT;
  Events ev;
  HIPCHK(c, ev.create((size_t)nsteps * 3 + 2));
  double adv = 0;
  int64_t chunks = 0;
  for (int64_t s0 = 0; s0 < nsteps; s0 += kChunk) {
    const int cnt = (int)std::min<int64_t>(kChunk, nsteps - s0);
    for (int s = 0; s < cnt; ++s) {
      hipError_t e = launch_step(c, c->rot + s, s, dt, fp32, &ev[(size_t)(s0 + s) * 3]);
      if (e != hipSuccess) return fail(c, SEM_E_HIP, std::string("launch: ") + hipGetErrorString(e));
    }
    HIPCHK(c, hipEventRecord(ev[ev.size() - 2], c->stream));
    hipError_t e = launch_advance(c, cnt, nn_first, nsteps);
    if (e != hipSuccess) return fail(c, SEM_E_HIP, std::string("launch: ") + hipGetErrorString(e));
    HIPCHK(c, hipEventRecord(ev[ev.size() - 1], c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    float ms = 0;
    HIPCHK(c, hipEventElapsedTime(&ms, ev[ev.size() - 2], ev[ev.size() - 1]));
    adv += ms;
    ++chunks;
    c->rot = (int)((c->rot + cnt) % 3);
  }
  double acc[2] = {0, 0};
  for (int64_t s = 0; s < nsteps; ++s)
    for (int k = 0; k < 2; ++k) {
      float ms = 0;
      HIPCHK(c, hipEventElapsedTime(&ms, ev[(size_t)s * 3 + k], ev[(size_t)s * 3 + k + 1]));
      acc[k] += ms;
    }
  ms_out[0] = acc[0] / nsteps;
  ms_out[1] = acc[1] / nsteps;
  ms_out[2] = adv / nsteps;
  return SEM_OK;
}

int sem_apply_K(sem_ctx* c, const double* u, double* f, int precision) {
  if (!c || !u || !f) return SEM_E_ARG;
  if (!c->kind || !c->have_material) return fail(c, SEM_E_STATE, "mesh and material must be set");
  if (precision != SEM_FP64 && precision != SEM_FP32_SPLIT) return fail(c, SEM_E_UNSUPPORTED, "unknown precision");
  if (const int rc = kernel_ready(c, precision == SEM_FP32_SPLIT); rc != SEM_OK) return rc;
  HIPCHK(c, hipSetDevice(c->device));
  const size_t b = c->dof * sizeof(double);
  double *du = nullptr, *df = nullptr;
  HIPCHK(c, hipMalloc((void**)&du, b));
  hipError_t e = hipMalloc((void**)&df, b);
  if (e != hipSuccess) { hipFree(du); return fail(c, SEM_E_HIP, "alloc"); }
  const bool fp32 = precision == SEM_FP32_SPLIT;
  std::vector<double> h(c->dof);
  to_internal(c, u, h.data());
  e = hipMemcpyAsync(du, h.data(), b, hipMemcpyHostToDevice, c->stream);
  // nodes that belong to no element get no force (the kernels write only the others)
  if (e == hipSuccess) e = hipMemsetAsync(df, 0, b, c->stream);
  const semk::UpdateArgs ua = update_args(c, nullptr, du, df, 0.0, 0);
  if (e == hipSuccess)
    e = semk::launch_patch(c->kind, c->nxy, c->nz, fp32, c->affine, 1, c->variant, c->plan.P, c->plan.n_patch, c->d_gtab,
                           patch_args(c, nullptr), ua, c->stream);
  if (e == hipSuccess) e = semk::launch_shared(c->nf, true, fp32, shared_args(c), ua, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(h.data(), df, b, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e == hipSuccess) to_external(c, h.data(), f);
  hipFree(du);
  hipFree(df);
  if (e != hipSuccess) return fail(c, SEM_E_HIP, std::string("apply_K: ") + hipGetErrorString(e));
  return SEM_OK;
}

int sem_apply_K_noalloc(sem_ctx* c, const double* u, double* f, int precision) {
  if (!c || !u || !f) return SEM_E_ARG;
  if (!c->kind || !c->have_material) return fail(c, SEM_E_STATE, "mesh and material must be set");
  if (precision != SEM_FP64 && precision != SEM_FP32_SPLIT) return fail(c, SEM_E_UNSUPPORTED, "unknown precision");
  if (const int rc = kernel_ready(c, precision == SEM_FP32_SPLIT); rc != SEM_OK) return rc;
  if (!c->scratch_dev) return fail(c, SEM_E_HIP, "the allocation-free scratch is not available on this device");
  // scratch layout (doubles): internal u [dof] | internal f [dof] | partials [n_slots x nf]
  const size_t need = (2 * (size_t)c->dof + (size_t)c->plan.n_slots * c->nf) * sizeof(double);
  if (need > kScratchBytes)
    return fail(c, SEM_E_ARG, "mesh too large for sem_apply_K_noalloc: it needs " + std::to_string(need) +
                                  " bytes of scratch, the limit is " + std::to_string(kScratchBytes));
  HIPCHK(c, hipSetDevice(c->device));
  double* du = c->scratch_dev;
  double* df = du + c->dof;
  const int nf = c->nf;
  const int64_t n = c->n_nodes;
  const unsigned blocks = (unsigned)((n * nf + 255) / 256);
  hipLaunchKernelGGL(permute_in_kernel, dim3(blocks), dim3(256), 0, c->stream, u, du, c->d_node_new, n, nf);
  hipError_t e = hipGetLastError();
  semk::UpdateArgs ua = update_args(c, nullptr, du, df, 0.0, 0);
  ua.partial = df + c->dof;
  const bool fp32 = precision == SEM_FP32_SPLIT;
  if (e == hipSuccess)
    e = semk::launch_patch(c->kind, c->nxy, c->nz, fp32, c->affine, 1, c->variant, c->plan.P, c->plan.n_patch,
                           c->d_gtab, patch_args(c, nullptr), ua, c->stream);
  if (e == hipSuccess) e = semk::launch_shared(c->nf, true, fp32, shared_args(c), ua, c->stream);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(permute_out_kernel, dim3(blocks), dim3(256), 0, c->stream, df, f, c->d_node_new, n, nf,
                       c->n_used);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return fail(c, SEM_E_HIP, std::string("apply_K_noalloc: ") + hipGetErrorString(e));
  return SEM_OK;
}

int sem_adjoint(sem_ctx* c, double dt, int64_t nn_first, int64_t nn_last, int precision, const double* dJ,
                int64_t ld, double* g_stiff, double* g_rho, int64_t seg_len, double* seconds) {
  if (!c) return SEM_E_ARG;
  if (!c->kind || !c->have_material) return fail(c, SEM_E_STATE, "mesh and material must be set");
  if (!(dt > 0) || nn_first < 1 || nn_last < nn_first) return fail(c, SEM_E_ARG, "bad step range or dt");
  if (precision != SEM_FP64 && precision != SEM_FP32_SPLIT) return fail(c, SEM_E_UNSUPPORTED, "unknown precision");
  if (c->n_out > 0 && (!dJ || ld < nn_last)) return fail(c, SEM_E_ARG, "dJ must be [ld >= nn_last] x n_out");
  if (c->ex.ptr && nn_last > c->ex.nft) return fail(c, SEM_E_ARG, "signal shorter than the last step");
  if (nn_last + 2 > INT32_MAX) return fail(c, SEM_E_ARG, "step range too long");
  if (const int rc = kernel_ready(c, precision == SEM_FP32_SPLIT); rc != SEM_OK) return rc;
  HIPCHK(c, hipSetDevice(c->device));
  const bool fp32 = precision == SEM_FP32_SPLIT;
  const int64_t nsteps = nn_last - nn_first + 1;
  const int64_t S = seg_len > 0 ? std::min(seg_len, nsteps) : (int64_t)std::ceil(std::sqrt((double)nsteps));
  const int64_t nseg = (nsteps + S - 1) / S;
  const size_t db = c->dof * sizeof(double);
  const double a0 = 1.0 / (dt * dt), a1 = 1.0 / (2.0 * dt);
  // Adjoint sources.  The trace of output j at step nn is sum_t w_t q(u)[dof_t]; q reads
  // u_nn (disp), u_{nn+1} (disp after the update), a1 (u_{nn+1} - u_{nn-1}) (velocity) or
  // a0 (u_{nn+1} - 2 u_nn + u_{nn-1}) (acceleration).  g_k = dJ/du_k for the states
  // k = nn_first+1..nn_last+1 that the run determines, as signal j at column k-1, over dt^2
  // (the step kernel multiplies its forces by c1 = dt^2 / (1 + dt dv / 2)).
  SourceSet adj;
  const int n_sig = std::max(c->n_out, 1);
  const int64_t nft = nn_last + 1;
  {
    auto d = [&](int j, int64_t nn) { return (nn >= nn_first && nn <= nn_last) ? dJ[j * ld + nn - 1] : 0.0; };
    std::vector<double> st((size_t)n_sig * nft, 0.0);
    for (int j = 0; j < c->n_out; ++j)
      for (int64_t k = nn_first + 1; k <= nn_last + 1; ++k) {
        double g = 0.0;
        switch (c->out_q) {
          case SEM_Q_DISP: g = d(j, k); break;
          case SEM_Q_VEL: g = a1 * (d(j, k - 1) - d(j, k + 1)); break;
          case SEM_Q_DISP_POST: g = d(j, k - 1); break;
          default: g = a0 * (d(j, k - 1) - 2.0 * d(j, k) + d(j, k + 1)); break;
        }
        st[(size_t)j * nft + k - 1] = g / (dt * dt);
      }
    std::vector<int32_t> sig((size_t)c->n_terms);
    for (int j = 0; j < c->n_out; ++j)
      for (int64_t t = c->h_out_ptr[j]; t < c->h_out_ptr[j + 1]; ++t) sig[t] = j;
    const int rc = upload_sources(c, c->n_terms, c->h_out_dof.data(), c->h_out_w.data(), sig.data(), n_sig, nft,
                                  st.data(), adj);
    if (rc != SEM_OK) { adj.release(); return rc; }
  }
  // work buffers: checkpoints (u_{n0-1}, u_{n0}) of every segment, the recomputed states of
  // one segment, three adjoint states, the gradients (patch order)
  double *ck = nullptr, *seg = nullptr, *lam = nullptr, *gs = nullptr, *gr = nullptr;
  auto cleanup = [&]() {
    for (double* p : {ck, seg, lam, gs, gr})
      if (p) hipFree(p);
    adj.release();
  };
  hipError_t e = hipMalloc((void**)&ck, 2 * db * (size_t)nseg);
  if (e == hipSuccess) e = hipMalloc((void**)&seg, db * (size_t)S);
  if (e == hipSuccess) e = hipMalloc((void**)&lam, 3 * db);
  if (e == hipSuccess) e = hipMalloc((void**)&gs, c->n_el * sizeof(double));
  if (e == hipSuccess) e = hipMalloc((void**)&gr, c->n_el * sizeof(double));
  if (e != hipSuccess) {
    cleanup();
    return fail(c, SEM_E_NOMEM, std::string("adjoint work buffers: ") + hipGetErrorString(e));
  }
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  const uint64_t ctr[semk::kCtr] = {0, UINT64_MAX, 0, 0};
  e = hipEventCreate(&ev0);
  if (e == hipSuccess) e = hipEventCreate(&ev1);
  if (e == hipSuccess) e = hipEventRecord(ev0, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(c->d_ctr, ctr, sizeof(ctr), hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) e = hipMemsetAsync(lam, 0, 3 * db, c->stream);
  if (e == hipSuccess) e = hipMemsetAsync(gs, 0, c->n_el * sizeof(double), c->stream);
  if (e == hipSuccess) e = hipMemsetAsync(gr, 0, c->n_el * sizeof(double), c->stream);
  // forward sweep, keeping the first two states of every segment
  int r = c->rot;
  for (int64_t i = 0; i < nseg && e == hipSuccess; ++i) {
    const int64_t n0 = nn_first + i * S, n1 = std::min(n0 + S - 1, nn_last);
    e = hipMemcpyAsync(ck + (size_t)(2 * i) * c->dof, c->d_u[r % 3], db, hipMemcpyDeviceToDevice, c->stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(ck + (size_t)(2 * i + 1) * c->dof, c->d_u[(r + 1) % 3], db, hipMemcpyDeviceToDevice,
                         c->stream);
    for (int64_t nn = n0; nn <= n1 && e == hipSuccess; ++nn, ++r)
      e = launch_plain_step(c, c->d_u[r % 3], c->d_u[(r + 1) % 3], c->d_u[(r + 2) % 3], nn, dt, fp32, &c->ex);
  }
  // reverse sweep: lambda_n lives in lam[(nn_last + 2 - n) % 3]; lambda_{nn_last+1} and
  // lambda_{nn_last+2} are zero.  Segment slot m holds u_{n0-1+m}: slots 0, 1 are the
  // checkpoint, slots 2.. the recomputed states.
  auto L = [&](int64_t n) { return lam + (size_t)((nn_last + 2 - n) % 3) * c->dof; };
  for (int64_t i = nseg - 1; i >= 0 && e == hipSuccess; --i) {
    const int64_t n0 = nn_first + i * S, n1 = std::min(n0 + S - 1, nn_last);
    double* ck0 = ck + (size_t)(2 * i) * c->dof;
    auto slot = [&](int64_t m) { return m < 2 ? ck0 + (size_t)m * c->dof : seg + (size_t)(m - 2) * c->dof; };
    for (int64_t nn = n0; nn <= n1 && e == hipSuccess; ++nn) {
      const int64_t m = nn - n0 + 1;
      e = launch_plain_step(c, slot(m - 1), slot(m), slot(m + 1), nn, dt, fp32, &c->ex);
    }
    for (int64_t n = n1; n >= n0 && e == hipSuccess; --n) {
      // (M + dt/2 C) lambda_n = g_{n+1} + (2M - dt^2 K) lambda_{n+1} - (M - dt/2 C) lambda_{n+2}:
      // the step kernel with U = lambda_{n+1}, U_old = lambda_{n+2}, sources at step n+1
      e = launch_plain_step(c, L(n + 2), L(n + 1), L(n), n + 1, dt, fp32, &adj);
      const int64_t m = n - n0 + 1;
      if (e == hipSuccess) e = launch_gradient(c, L(n), slot(m - 1), slot(m), slot(m + 1), dt, fp32, gs, gr);
    }
  }
  if (e == hipSuccess) e = hipEventRecord(ev1, c->stream);
  if (e == hipSuccess) e = hipEventSynchronize(ev1);
  float ms = 0;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, ev0, ev1);
  std::vector<double> hs(c->n_el), hr(c->n_el);
  if (e == hipSuccess) e = hipMemcpy(hs.data(), gs, c->n_el * sizeof(double), hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(hr.data(), gr, c->n_el * sizeof(double), hipMemcpyDeviceToHost);
  if (ev0) hipEventDestroy(ev0);
  if (ev1) hipEventDestroy(ev1);
  cleanup();
  if (e != hipSuccess) return fail(c, SEM_E_HIP, std::string("adjoint: ") + hipGetErrorString(e));
  c->rot = (int)((c->rot + nsteps) % 3);
  for (int64_t k = 0; k < c->n_el; ++k) {
    if (g_stiff) g_stiff[c->plan.perm[k]] = hs[k];
    if (g_rho) g_rho[c->plan.perm[k]] = hr[k];
  }
  if (seconds) *seconds = ms * 1e-3;
  return SEM_OK;
}

}  // extern "C"
