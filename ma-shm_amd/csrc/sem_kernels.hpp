// Device kernels of the spectral-element time step (gfx950).
//
// One step = three launches on the context's stream:
//   1. *_forces  : element kernel.  Gathers the element's nodal values through the
//                  connectivity, forms the GLL derivatives by sum factorisation in LDS,
//                  applies the constitutive law and the transposed derivatives, adds the
//                  point excitation, and writes element-local forces f_loc[e][p][field].
//   2. node_update: node kernel.  Sums each node's element-local forces through the
//                  node->(element,point) inverse map in a fixed order (the reference's
//                  basket order), applies the central-difference update with the lumped
//                  inverse mass and writes U^{n+1}; at frame steps also the frame and the
//                  divergence guard.  No float atomics: runs are bitwise reproducible.
//   3. post_step : one block.  Linear output functionals (sensor traces) and the step
//                  counters (the step index lives on the device so that a captured graph
//                  of steps can be replayed).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace semk {

constexpr int kMaxN = 10;

// GLL tables live in a small device buffer `gtab`:
//   [ D (n*n) | Dz (nz*nz, solid only) | w (n) | wz (nz, solid only) ],
// D[i*n+j] = L_j'(x_i) (the reference's Sxprim, shape2D_prim.m:41-49).

struct ForceArgs {
  int64_t n_el;
  const int32_t* nodes;     // [n_el][n_loc]
  const double* U;          // [dof]
  const double* geom;       // affine: [n_el][G]; else [n_el][n_loc][G]
  const int32_t* elem_mat;  // [n_el]
  const double* mat;        // [n_mat][n_coef]
  int n_coef;
  const double* sscale;     // [n_el] stiffness scale
  void* floc;               // [n_el][n_loc][nf] of T
  // excitation (element-local CSR): entries ex_ptr[e]..ex_ptr[e+1]
  const int32_t* ex_ptr;    // [n_el+1] or null
  const int32_t* ex_pf;     // point*nf + field
  const double* ex_amp;
  const double* st;         // signal, read at st[nn-1]
  const int64_t* ctr;       // ctr[0] = nn
  int excite;               // 0: pure operator
};

struct NodeArgs {
  int64_t n_nodes;
  const int32_t* rowptr;
  const int32_t* occ;
  const void* floc;
  const double* Uold;
  const double* U;
  double* Unew;          // update: U^{n+1}; apply-K: F_int
  const double* minv;    // [dof]
  double c1[5], c2[5], c3[5];
  double a0, a1;
  // frames / guard
  int n_frames;
  const int64_t* frame_nn;
  int frame_q;
  double* frames;        // [n_frames][dof]
  int64_t* ctr;          // [0]=nn [1]=frame index [2]=guard flag [3]=diverged nn
  double guard_limit;    // <= 0: off
  int guard_field;       // -1 all
  int guard_post;
};

struct PostArgs {
  int n_out;
  const int64_t* out_ptr;
  const int64_t* out_dof;
  const double* out_w;
  int out_q;
  const double* Uold;
  const double* U;
  const double* Unew;
  double a0, a1;
  double* traces;       // column-major [ld][n_out]
  int64_t ld;
  int64_t nn_base;      // traces row = nn - nn_base
  int n_frames;
  const int64_t* frame_nn;
  int64_t* ctr;
};

// launch helpers (defined in sem_kernels.hip)
hipError_t launch_shell_forces_tab(int n, bool fp32, bool affine, const double* gtab,
                                   const ForceArgs& a, hipStream_t s);
hipError_t launch_solid_forces_tab(int n, int nz, bool fp32, bool affine, const double* gtab,
                                   const ForceArgs& a, hipStream_t s);
hipError_t launch_node_update(int nf, bool fp32, bool applyK, const NodeArgs& a, hipStream_t s);
hipError_t launch_post_step(const PostArgs& a, hipStream_t s);
bool shell_supported(int n);
bool solid_supported(int n, int nz);

}  // namespace semk
