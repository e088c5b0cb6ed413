// Wavefield on a uniform grid (spec2meshgrid_flat_shell.m:1-583) for gfx950.
//
// Set-up on the host (semhost::grid_plan): owner element and local coordinates of every
// grid point, kept on the device as the owner's node ids [P][n^2] and the 1D Lagrange
// weights [P][n] in each direction.  Per frame, one thread per grid point gathers the
// owner's n^2 nodal values of the requested quantity and sums N_k U_k with
// N_k = L_i(s) L_j(t), k = j n + i (the reference's Ns*ZA, one sparse row per point).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/sem_hip.h"
#include "sem_host.hpp"

struct sem_grid {
  int device = 0;
  hipStream_t stream = nullptr;
  semhost::GridPlan plan;
  int64_t n_nodes = 0;
  int32_t* d_pn = nullptr;  // [P][n^2] owner node ids
  double* d_Lx = nullptr;   // [P][n]
  double* d_Ly = nullptr;   // [P][n]
  std::string err;
};

namespace {

// quantity `motion` (1-9) of node nd in one shell frame (u, phix, v, phiy, w per node),
// hs = +h/2 (top surface) or -h/2 (bottom): spec2meshgrid_flat_shell.m:321-565
__device__ __forceinline__ double quantity(const double* __restrict__ U, int64_t nd, int motion, double hs) {
  const double* u = U + nd * 5;
  switch (motion) {
    case 1: return u[0];
    case 2: return u[2];
    case 3: return u[4];
    case 4: return u[1];
    case 5: return u[3];
    case 6: return u[0] + hs * u[1];
    case 7: return u[2] + hs * u[3];
    case 8: {
      const double a = u[0] + hs * u[1], b = u[2] + hs * u[3];
      return sqrt(a * a + b * b);
    }
    default: {
      const double a = u[0] + hs * u[1], b = u[2] + hs * u[3];
      return sqrt(a * a + b * b + u[4] * u[4]);
    }
  }
}

template <int N>
__global__ __launch_bounds__(256) void grid_eval_kernel(const int32_t* __restrict__ pn, const double* __restrict__ Lx,
                                                        const double* __restrict__ Ly, const int32_t* __restrict__ remap,
                                                        const double* __restrict__ frames, int64_t dof, int64_t P,
                                                        int Ny, int Nx, int motion, double hs, double* __restrict__ out) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  const double* U = frames + (int64_t)blockIdx.y * dof;
  double lx[N], ly[N];
#pragma unroll
  for (int i = 0; i < N; ++i) {
    lx[i] = Lx[p * N + i];
    ly[i] = Ly[p * N + i];
  }
  const int32_t* ids = pn + p * N * N;
  double acc = 0.0;
#pragma unroll
  for (int j = 0; j < N; ++j)
#pragma unroll
    for (int i = 0; i < N; ++i) {
      int64_t nd = ids[j * N + i];
      if (remap) nd = remap[nd];
      acc += (lx[i] * ly[j]) * quantity(U, nd, motion, hs);
    }
  const int64_t c = p / Ny, r = p - c * Ny;  // point p = r + Ny*c at (xi[c], yi[r])
  out[(int64_t)blockIdx.y * P + r * Nx + c] = acc;
}

template <typename T>
hipError_t upload(T*& d, const T* h, size_t n, hipStream_t s) {
  if (d) { hipFree(d); d = nullptr; }
  hipError_t e = hipMalloc((void**)&d, std::max<size_t>(n, 1) * sizeof(T));
  if (e == hipSuccess && n) e = hipMemcpyAsync(d, h, n * sizeof(T), hipMemcpyHostToDevice, s);
  return e;
}

thread_local std::string create_err;  // why the last sem_grid_create on this thread failed

int fail(sem_grid* g, int code, const std::string& msg) {
  if (g) g->err = msg;
  return code;
}

int hip_fail(sem_grid* g, hipError_t e) { return fail(g, SEM_E_HIP, hipGetErrorString(e)); }

}  // namespace

extern "C" {

int sem_grid_create(sem_grid** out, int device, int64_t n_el, int n_xy, const int32_t* nodes, int64_t n_nodes,
                    const double* coords, int64_t n_search, int Nx, int Ny) {
  if (!out) return SEM_E_ARG;
  *out = nullptr;
  if (n_el <= 0 || n_xy < 2 || n_xy > 8 || !nodes || n_nodes <= 0 || !coords || Nx < 2 || Ny < 2)
    return SEM_E_ARG;
  const int nl = n_xy * n_xy;
  for (int64_t q = 0; q < n_el * nl; ++q)
    if (nodes[q] < 0 || nodes[q] >= n_nodes) return SEM_E_ARG;
  sem_grid* g = new sem_grid();
  g->device = device;
  g->n_nodes = n_nodes;
  std::string err;
  if (!semhost::grid_plan(n_el, n_xy, nodes, n_nodes, coords, n_search > 0 ? n_search : n_el, Nx, Ny, g->plan,
                          err)) {
    create_err = err;
    delete g;
    return SEM_E_ARG;
  }
  hipError_t e = hipSetDevice(device);
  if (e == hipSuccess) e = hipStreamCreate(&g->stream);
  const int64_t P = (int64_t)Nx * Ny;
  std::vector<int32_t> pn((size_t)P * nl);
  for (int64_t j = 0; j < P; ++j)
    std::memcpy(&pn[(size_t)j * nl], nodes + (int64_t)g->plan.owner[j] * nl, nl * sizeof(int32_t));
  if (e == hipSuccess) e = upload(g->d_pn, pn.data(), pn.size(), g->stream);
  if (e == hipSuccess) e = upload(g->d_Lx, g->plan.Lx.data(), g->plan.Lx.size(), g->stream);
  if (e == hipSuccess) e = upload(g->d_Ly, g->plan.Ly.data(), g->plan.Ly.size(), g->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(g->stream);
  if (e != hipSuccess) {
    create_err = hipGetErrorString(e);
    sem_grid_destroy(g);
    return SEM_E_HIP;
  }
  *out = g;
  return SEM_OK;
}

int sem_grid_destroy(sem_grid* g) {
  if (!g) return SEM_OK;
  hipSetDevice(g->device);
  if (g->d_pn) hipFree(g->d_pn);
  if (g->d_Lx) hipFree(g->d_Lx);
  if (g->d_Ly) hipFree(g->d_Ly);
  if (g->stream) hipStreamDestroy(g->stream);
  delete g;
  return SEM_OK;
}

const char* sem_grid_last_error(const sem_grid* g) { return g ? g->err.c_str() : create_err.c_str(); }

int sem_grid_points(const sem_grid* g, double* xi, double* yi, int32_t* owner, double* s, double* t) {
  if (!g) return SEM_E_ARG;
  const auto& p = g->plan;
  if (xi) std::copy(p.xi.begin(), p.xi.end(), xi);
  if (yi) std::copy(p.yi.begin(), p.yi.end(), yi);
  if (owner) std::copy(p.owner.begin(), p.owner.end(), owner);
  if (s) std::copy(p.s.begin(), p.s.end(), s);
  if (t) std::copy(p.t.begin(), p.t.end(), t);
  return SEM_OK;
}

int sem_grid_eval(sem_grid* g, int64_t n_frames, int64_t dof, const double* frames, int motion, int surface,
                  double h, int64_t n_twin, const int32_t* twin_under, const int32_t* twin_above, double* data) {
  if (!g || n_frames <= 0 || !frames || !data || motion < 1 || motion > 9 || surface == 0 || n_twin < 0 ||
      (n_twin && (!twin_under || !twin_above)))
    return fail(g, SEM_E_ARG, "invalid argument");
  if (dof != 5 * g->n_nodes) return fail(g, SEM_E_ARG, "frames must hold 5 DOF per mesh node");
  if (n_frames > 65535) return fail(g, SEM_E_ARG, "at most 65535 frames per call");
  hipError_t e = hipSetDevice(g->device);
  if (e != hipSuccess) return hip_fail(g, e);
  const int64_t P = (int64_t)g->plan.Nx * g->plan.Ny;
  double* d_frames = nullptr;
  double* d_out = nullptr;
  int32_t* d_remap = nullptr;
  // top surface: the nodes under a delamination take the values of their twins above it
  // (:567-569), as a node remap applied at the gather
  if (surface > 0 && n_twin > 0) {
    std::vector<int32_t> remap(g->n_nodes);
    for (int64_t k = 0; k < g->n_nodes; ++k) remap[k] = (int32_t)k;
    for (int64_t q = 0; q < n_twin; ++q) {
      if (twin_under[q] < 0 || twin_under[q] >= g->n_nodes || twin_above[q] < 0 || twin_above[q] >= g->n_nodes)
        return fail(g, SEM_E_ARG, "twin node id out of range");
      remap[twin_under[q]] = twin_above[q];
    }
    e = upload(d_remap, remap.data(), remap.size(), g->stream);
  }
  if (e == hipSuccess) e = upload(d_frames, frames, (size_t)(n_frames * dof), g->stream);
  if (e == hipSuccess) e = hipMalloc((void**)&d_out, (size_t)(n_frames * P) * sizeof(double));
  if (e == hipSuccess) {
    const dim3 grid((unsigned)((P + 255) / 256), (unsigned)n_frames);
    const double hs = surface > 0 ? h / 2 : -h / 2;
#define GO(NV)                                                                                                 \
  hipLaunchKernelGGL(grid_eval_kernel<NV>, grid, dim3(256), 0, g->stream, g->d_pn, g->d_Lx, g->d_Ly, d_remap, \
                     d_frames, dof, P, g->plan.Ny, g->plan.Nx, motion, hs, d_out)
    switch (g->plan.n) {
      case 2: GO(2); break;
      case 3: GO(3); break;
      case 4: GO(4); break;
      case 5: GO(5); break;
      case 6: GO(6); break;
      case 7: GO(7); break;
      default: GO(8); break;
    }
#undef GO
    e = hipGetLastError();
  }
  if (e == hipSuccess)
    e = hipMemcpyAsync(data, d_out, (size_t)(n_frames * P) * sizeof(double), hipMemcpyDeviceToHost, g->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(g->stream);
  if (d_frames) hipFree(d_frames);
  if (d_out) hipFree(d_out);
  if (d_remap) hipFree(d_remap);
  return e == hipSuccess ? SEM_OK : hip_fail(g, e);
}

}  // extern "C"
