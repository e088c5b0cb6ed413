// Device kernels of the spectral-element time step (gfx950).
//
// Elements are grouped into patches (a few neighbouring elements, e.g. 2x2x2 hexahedra
// or 4x4 shells).  One step = two launches on the context's stream:
//   1. *_patch_kernel (one workgroup per patch): gathers the patch's nodal values once
//      into LDS, forms the GLL derivatives of every element by sum factorisation in LDS,
//      applies the constitutive law and the transposed derivatives, adds the point
//      excitation, and sums the element-local forces of every patch node in LDS in a
//      fixed order (the reference's basket order).  Nodes whose elements all lie in the
//      patch ("owned") get the central-difference update right there; nodes on the
//      patch boundary write their partial sum to a slot of `partial`.
//   2. shared_update_kernel: for every patch-boundary node, sums its partials in a fixed
//      order and applies the update.
// Both kernels also store the sampled quantity of every sensor term of their nodes in a
// ring buffer (one row per step of a chunk); once per chunk of steps output_kernel forms
// the output functionals from the ring (one wavefront per step and output) and
// counter_kernel advances the step counter.  Kernels read the step as ctr[0] + step offset, so a captured graph of a chunk
// of steps can be replayed.  No float atomics and no device-scope fences in the step:
// runs are bitwise reproducible.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace semk {

// GLL tables live in a small device buffer `gtab`:
//   [ D (n*n) | Dz (nz*nz, solid only) | w (n) | wz (nz, solid only) ] as double, then the
// same values as float, D[i*n+j] = L_j'(x_i) (the reference's Sxprim, shape2D_prim.m:41-49).

// Device counters: [0] first step of the current chunk, [1] earliest step at which the
// divergence guard tripped (UINT64_MAX = never).
constexpr int kCtr = 4;

// The lumped mass of a node has one coefficient per kind of field: a solid's three
// displacements share w·detJ·rho (mass3D_v2.m:51-57), a shell's translations share m11 and
// its two rotations I11 (force_excit_c.m:497-503: [m11, I11, m11, I11, m11]).  The inverse
// mass is stored per node with one column per coefficient; field f reads column
// mass_col(nf, f).
__host__ __device__ constexpr int mass_cols(int nf) { return nf == 5 ? 2 : 1; }
__host__ __device__ constexpr int mass_col(int nf, int f) { return nf == 5 ? (f & 1) : 0; }

struct UpdateArgs {
  const double* U;        // U^n
  const double* Uold;     // U^{n-1}
  double* Unew;           // U^{n+1}; apply-K: F_int
  const double* minv;     // [n_nodes][mass_cols(nf)] inverse lumped mass (both forms)
  double c1[5], c2[5], c3[5];
  double a0, a1;
  int step_off;           // step = ctr[0] + step_off
  const int32_t* frame_tab;  // [n_ftab] frame index of step nn at nn-1, or -1
  int64_t n_ftab;
  int frame_q;
  double* frames;         // [n_frames][dof]
  uint64_t* ctr;
  double guard_limit;     // <= 0: off
  int guard_field;        // -1: all fields
  int guard_post;
  int64_t nan_dof;        // DOF whose U^{n+1} is tested for NaN (the caller's DOF 2)
  int64_t dof;
  void* partial;          // [n_slots][nf], double (FP64 form) or float (FP32 form)
  // sensor terms
  int out_q;
  double* ring;           // [chunk][n_terms]
  int64_t n_terms;
  // gradient accumulation (adjoint): with U = u_n, Uold = u_{n-1}, Unew = u_{n+1} (read
  // only) and lam = lambda_n, per element e (patch order)
  //   gs[e] += gscale_s * lambda^T K_e^0 u_n
  //   gr[e] += gscale_m * lambda^T M_e^0 [(u_{n+1} - 2u_n + u_{n-1}) + gdv_f (u_{n+1} - u_{n-1})]
  const double* lam;
  double* gs;
  double* gr;
  double gscale_s, gscale_m;
  double gdv[5];
};

struct PatchArgs {
  const int32_t* el0;       // [n_patch+1] element range of each patch
  const int32_t* pn_base;   // [n_patch+1] range of each patch's nodes in pn_id
  const int32_t* pn_owned;  // [n_patch] owned nodes come first
  const int32_t* pn_id;     // patch nodes (internal ids); owned ones are pn_id[pn0] + ln
  const int32_t* slot_base; // [n_patch] partial slot of the patch's first boundary node (ln = owned)
  // a patch's local structure, stored once per distinct one (template): its pn+1 occurrence
  // offsets at occ_ptr + optr_base[b], its occurrence lists (LDS point index le*NP + p) at
  // occ + occ_base[b], and the patch-local node of each point of its element le at
  // lmap[lm_base[b] + le*NP + p]
  const uint16_t* occ_ptr;
  const int32_t* optr_base; // [n_patch]
  const int32_t* occ_base;  // [n_patch]
  const uint16_t* occ;
  // the template's occurrence records, [pn][8] at occ8 + o8_base[b] (semhost::PatchPlan::occ8)
  const uint16_t* occ8;
  const int32_t* o8_base;   // [n_patch]
  const int32_t* lm_base;   // [n_patch]
  const uint16_t* lmap;
  int max_pn;               // LDS rows reserved for patch nodes
  const double* geom;       // affine: [n_el][G]; else [n_el][NP][G]
  const float* geomf;       // non-affine: the per-point geometry as float (FP32 form)
  const int32_t* elem_mat;
  const double* mat;
  int n_coef;
  const double* sscale;     // nullable
  int diag;                 // affine solid with diagonal inverse Jacobians (staged kernel only)
  const int32_t* ex_ptr;    // [n_el+1] element-local excitation CSR (nullable)
  const int32_t* ex_pf;     // point*nf + field
  const int32_t* ex_sig;    // signal of each entry
  const double* ex_amp;
  const double* st;         // [n_sig][ex_nft], step nn reads column nn-1
  int64_t ex_nft;
  int excite;
  // sensor terms of owned nodes: entries tp_ptr[b]..tp_ptr[b+1] (ln*nf + f, term index)
  const int32_t* tp_ptr;    // nullable
  const int32_t* tp_lf;
  const int32_t* tp_t;
  const uint64_t* mtab;     // MFMA shell kernel: A-operand tables (nullable)
};

struct SharedArgs {
  int64_t n_shared;
  const int32_t* sh_id;     // [n_shared] internal id of each boundary node
  const int32_t* sh_ptr;    // [n_shared+1]: boundary node k has entries sh_ptr[k]..sh_ptr[k+1]
  const int32_t* sh_slot;   // the partial slot of each entry, in patch order
  const int32_t* ts_ptr;    // [n_shared+1] sensor terms (nullable)
  const int32_t* ts_f;
  const int32_t* ts_t;
};

struct AdvanceArgs {
  int count;                // steps in the finished chunk
  int n_out;
  const int64_t* out_ptr;   // [n_out+1] over terms
  const double* out_w;      // [n_terms]
  const double* ring;
  int64_t n_terms;
  double* traces;           // column-major [ld][n_out]
  int64_t ld;
  int64_t nn_base;          // trace row = nn - nn_base
  uint64_t* ctr;
};

// Element kernels ("variants"): 0 = point-wise contractions (shell: 16-element patches,
// solid: 4-element patches); 1 = line-wise contractions (shell: 8- or 16-element patches,
// solid: 4- or 8-element patches); 2 = contractions on the matrix cores (shell, FP32 form,
// 16-element patches, orders 4 and 5); 3 = line-wise contractions staged one direction at a
// time with per-point registers (solid, 4- or 8-element patches).  A plan is built for one patch size; the variant
// must accept it.
int shell_patch_elems(int n, bool fp32);
int solid_patch_elems(int n, int nz, bool fp32);
bool shell_supported(int n);
bool solid_supported(int n, int nz);
bool variant_accepts(int kind, int variant, int patch_elems);
// LDS bytes of a patch kernel for the given number of patch-node rows: all of it (dynamic
// and static), and the dynamic part a launch asks for.
size_t patch_lds_bytes(int kind, int n, int nz, bool fp32, int max_pn, int variant, int patch_elems);
size_t patch_dynamic_lds_bytes(int kind, int n, int nz, bool fp32, int max_pn, int variant, int patch_elems);
size_t solid_line_lds_bytes(int n, int nz, bool fp32, int max_pn, int patch_elems);
size_t solid_stage_lds_bytes(int n, int nz, bool fp32, int max_pn, int patch_elems);
size_t shell_line_lds_bytes(int n, bool fp32, int max_pn, int patch_elems);

// mode: 0 time step, 1 apply K (F_int into ua.Unew), 2 gradient accumulation
hipError_t launch_patch(int kind, int n, int nz, bool fp32, bool affine, int mode, int variant,
                        int patch_elems, int64_t n_patch, const double* gtab, const PatchArgs& pa,
                        const UpdateArgs& ua, hipStream_t s);
size_t shell_mfma_lds_bytes(int n, int max_pn);
size_t shell_mfma_table_entries(int n);
bool shell_mfma_supported(int n);
hipError_t launch_shell_mfma(int n, bool affine, int mode, int64_t n_patch, const double* gtab,
                             const PatchArgs& pa, const UpdateArgs& ua, hipStream_t s);
hipError_t launch_shell_line(int n, bool fp32, bool affine, int mode, int patch_elems, int64_t n_patch,
                             const double* gtab, const PatchArgs& pa, const UpdateArgs& ua, hipStream_t s);
hipError_t launch_solid_line(int n, int nz, bool fp32, bool affine, int mode, int patch_elems,
                             int64_t n_patch, const double* gtab, const PatchArgs& pa,
                             const UpdateArgs& ua, hipStream_t s);
hipError_t launch_solid_stage(int n, int nz, bool fp32, bool affine, int mode, int patch_elems,
                              int64_t n_patch, const double* gtab, const PatchArgs& pa,
                              const UpdateArgs& ua, hipStream_t s);
hipError_t launch_shared(int nf, bool applyK, bool fp32, const SharedArgs& sa, const UpdateArgs& ua,
                         hipStream_t s);
hipError_t launch_advance(const AdvanceArgs& a, hipStream_t s);

}  // namespace semk
