/*
 * sem_hip.h -- C-ABI of the MI355X spectral-element time-stepping library.
 *
 * Drop-in boundary for the solver loop of the reference's MATLAB engines
 * (pawelkudela/ma-shm).  The reference has no FFI; its operator boundary is the
 * engine signature plus the mode switch 'cpu'|'gpu':
 *   t_frames = main_flat_shell_force_excit_c(test_case,input_no,meshfile,mode,output_name,tasks)
 *     (src/models/flat_shell/flat_shell_engine/main_flat_shell_force_excit_c.m:1)
 *   3D script src/models/solid3D/main_spec3Dcuda_global_cpu_gpu3_wave_sens_opt_single_PCT_mat.m:1-4
 * A MATLAB caller reaches these entry points from a new mode='hip' branch through a
 * MEX gateway; Python reaches them through ctypes.
 *
 * Conventions: every call returns int (SEM_OK = 0, negative = error, message via
 * sem_last_error).  A refused setter leaves the context as it was (the mesh, material,
 * sources, outputs and frames set before stay in force); a refused mesh upload leaves it
 * without a mesh.  No exceptions cross the ABI.
 * The caller owns all host buffers;
 * the library copies inputs on upload and writes into caller-allocated outputs.
 * A context is not thread-safe but may move between threads; separate contexts may be
 * used from separate threads at the same time (each keeps its own stream, buffers and
 * error message); sem_run blocks.
 * Node ids are 0-based int32; a DOF id is nf*node + field (the reference's
 * 5*node-4.. / 3*node-2.. numbering shifted to 0-based).  Matrices are row-major
 * unless stated otherwise; traces are column-major like MATLAB's displ(nft,n_out).
 */
#ifndef SEM_HIP_H
#define SEM_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sem_ctx sem_ctx;

enum {
  SEM_OK = 0,
  SEM_E_ARG = -1,        /* invalid argument / shape */
  SEM_E_HIP = -2,        /* HIP runtime error */
  SEM_E_STATE = -3,      /* call out of order (e.g. run before mesh) */
  SEM_E_DIVERGED = -4,   /* divergence guard tripped; partial traces written */
  SEM_E_NOMEM = -5,
  SEM_E_UNSUPPORTED = -6
};

/* element kinds */
enum { SEM_SHELL5 = 1, /* flat shell, 5 DOF/node u,phix,v,phiy,w */
       SEM_SOLID3 = 2  /* hexahedral solid, 3 DOF/node u,v,w */ };

/* arithmetic forms of the element operator */
enum { SEM_FP64 = 0,       /* plain FP64 everywhere */
       SEM_FP32_SPLIT = 1  /* FP32 element arithmetic and local forces; FP64 state,
                              FP64 assembly sums and FP64 inverse mass */ };

/* recorded quantities */
enum { SEM_Q_DISP = 0,      /* U^n before the update (force_excit_c.m:701-705)          */
       SEM_Q_VEL = 1,       /* a1*(U^{n+1}-U^{n-1}) (main_flat_shell.m:774)             */
       SEM_Q_DISP_POST = 2, /* U^{n+1} (3D engine, main_spec3Dcuda_global_..._mat.m:1051) */
       SEM_Q_ACC = 3        /* a0*(U^{n-1}-2U^n+U^{n+1}) (force_excit_c.m:719)           */ };

/* --- context ------------------------------------------------------------ */
/* Replaces gpuDevice(1) / reset(g) (main_flat_shell.m:931-934). */
int sem_create(sem_ctx** ctx, int device);
int sem_destroy(sem_ctx* ctx);
const char* sem_last_error(const sem_ctx* ctx);
/* Library version string. */
const char* sem_version(void);

/* --- host-side mesh front end (no GPU needed) ------------------------- */
/* change_turn_quad.m:1-110 (called before quad2spectral, automesh_rectangle.m:57): every
 * bilinear quad (0-based corner ids, n_quads x 4, modified in place) whose Jacobian
 * determinant is not positive at a corner is renumbered 1,4,3,2 (clockwise -> counter-
 * clockwise).  pos: n_pos x 3 corner coordinates; *n_flipped (nullable) = quads renumbered. */
int sem_mesh_change_turn_quad(int64_t n_quads, int32_t* quads, int64_t n_pos, const double* pos,
                              int64_t* n_flipped);
/* quad2spectral_Fiborek.m:1-173: bilinear quads (0-based corner ids, n_quads x 4,
 * counter-clockwise) -> spectral elements of `order` (order+1 nodes per edge).
 * nodes_out must hold n_quads*(order+1)^2 int32; coords_out (n_max x 3) with
 * n_max = n_fem + n_quads*(4*(order-1) + (order-1)^2).  *n_nodes receives the count. */
int sem_mesh_quad2spectral(int64_t n_quads, const int32_t* quads, int64_t n_fem,
                           const double* pos, int order, int32_t* nodes_out,
                           double* coords_out, int64_t* n_nodes);
/* extrude_mesh.m:1-85 (incl. its hlay=dz update): 2D spectral mesh -> Nz element
 * layers with nz GLL nodes through each; h[Nz] layer thicknesses.
 * nodes3d_out: (Nz*n_el2d) x (n_xy*n_xy*nz); coords3d_out: n2d*(Nz*(nz-1)+1) x 3. */
int sem_mesh_extrude(int64_t n_el2d, int n_xy, const int32_t* nodes2d, int64_t n_nodes2d,
                     const double* coords2d, int nz, int Nz, const double* h,
                     int32_t* nodes3d_out, double* coords3d_out);

/* nodesMaps_Fiborek.m:1-27 / parallel_LG_nodes_Modified_Zb2.m (context-free):
 * 12-basket colouring of the element-local points.  I_G/I_L are 1-based,
 * column-major [rows x 12] with rows = ceil(n_el*n_loc/12), 0 = empty slot. */
int sem_mesh_baskets(int64_t n_el, int n_loc, const int32_t* nodes, int64_t n_nodes,
                     int32_t* I_G, int32_t* I_L);

/* quad2spectral_Fiborek.m:153-156 (its boundary_nodes output): sorted nodes on the element
 * edges owned by one element.  bnd must hold n_el*4*n_xy entries; *n_bnd receives the count. */
int sem_mesh_boundary_nodes(int64_t n_el, int n_xy, const int32_t* nodes, int32_t* bnd, int64_t* n_bnd);
/* split_delam_nodes_flat_shell.m:1-276: the n_den delaminated elements den[] are appended
 * again as elements n_el..n_el+n_den-1 (den_above; den_under = den) and the copies get new,
 * coincident nodes except on the delamination's rim.  bnd: the plate's boundary nodes
 * (sem_mesh_boundary_nodes).  nodes_out holds (n_el+n_den) x n_xy^2, coords_out
 * (n_nodes + n_den*n_xy^2) x 3; *n_nodes_out receives the new node count. */
int sem_mesh_split_delam(int64_t n_el, int n_xy, const int32_t* nodes, int64_t n_nodes,
                         const double* coords, int64_t n_den, const int32_t* den, int64_t n_bnd,
                         const int32_t* bnd, int32_t* nodes_out, double* coords_out,
                         int64_t* n_nodes_out, int32_t* den_under, int32_t* den_above);

/* --- wavefield on a uniform grid (spec2meshgrid_flat_shell.m:1-583) ------- */
typedef struct sem_grid sem_grid;
/* Nx x Ny grid over the flat mesh's bounding box (1e-8 m inset; point j = r + Ny*c at
 * (xi[c], yi[r])), owner element among the first n_search elements (0 = all; the reference
 * leaves out the delaminated copies), local coordinates by the reference's Newton
 * iteration (:103-299).  The interpolation data stay on `device`.  On failure *g is NULL
 * and sem_grid_last_error(NULL) says why. */
int sem_grid_create(sem_grid** g, int device, int64_t n_el, int n_xy, const int32_t* nodes,
                    int64_t n_nodes, const double* coords, int64_t n_search, int Nx, int Ny);
int sem_grid_destroy(sem_grid* g);
const char* sem_grid_last_error(const sem_grid* g);
/* The same grid set-up on the host only (no device): grid coordinates xi[Nx], yi[Ny] and
 * per point (j = r + Ny*c) the owner element and its local xi/eta (outputs nullable). */
int sem_grid_plan(int64_t n_el, int n_xy, const int32_t* nodes, int64_t n_nodes, const double* coords,
                  int64_t n_search, int Nx, int Ny, double* xi, double* yi, int32_t* owner, double* s,
                  double* t);
/* grid coordinates xi[Nx], yi[Ny] and per point owner element, local xi/eta (nullable). */
int sem_grid_points(const sem_grid* g, double* xi, double* yi, int32_t* owner, double* s, double* t);
/* :321-575: quantity `motion` 1-9 (Ux Uy Uz Fix Fiy, Ux+-h/2 Fix, Uy+-h/2 Fiy, in-plane and
 * total magnitude) at the top (surface > 0) or bottom surface of n_frames shell frames
 * [n_frames x dof] (caller's DOF order, 5 per node).  Top surface: node twin_under[q] takes
 * the value of twin_above[q] (the delamination's split nodes, :567-569).  data receives
 * [n_frames x Ny x Nx], data[f][r][c] at (xi[c], yi[r]). */
int sem_grid_eval(sem_grid* g, int64_t n_frames, int64_t dof, const double* frames, int motion,
                  int surface, double h, int64_t n_twin, const int32_t* twin_under,
                  const int32_t* twin_above, double* data);

/* --- host-side material helpers ----------------------------------------- */
/* composite_plate_c_delam.m:10-19,48-87 (undelaminated element): lay plies, ply 1
 * on top; q is lay x 6 (q11 q12 q22 q44 q55 q66).  coef_out[23] =
 * a11 a12 a16 a22 a26 a66 a44 a45 a55 b11 b12 b16 b22 b26 b66 d11 d12 d16 d22 d26 d66 m11 I11 */
int sem_laminate_shell(int lay, const double* lh, const double* alpha_deg,
                       const double* rho, const double* q, double* coef_out);
/* composite_plate_c_delam.m:89-173: the sub-laminate of the elements above (side > 0:
 * plies 1..delamination_layer) or under (side < 0: the other plies) a delamination, at
 * their z in the whole laminate; same row layout as sem_laminate_shell. */
int sem_laminate_shell_delam(int lay, const double* lh, const double* alpha_deg,
                             const double* rho, const double* q, int delamination_layer, int side,
                             double* coef_out);
/* composite3D_transf.m:49-108 'smeared' with q given per ply (lay x 9: q11 q12 q13 q22
 * q23 q33 q44 q55 q66).  coef_out[14] = D11 D12 D13 D14 D22 D23 D24 D33 D34 D44 D55 D56
 * D66 rho, with the reference's pairing D55=a55 (gamma_yz), D66=a44 (gamma_xz). */
int sem_laminate_solid_smeared(int lay, const double* lh, const double* alpha_deg,
                               const double* rho, const double* q, double* coef_out);
/* composite3D_transf.m:109-119 'layered': one row per ply, coef_out[lay][14] in the layout of
 * sem_laminate_solid_smeared (the ply's rotated stiffness, same D55/D56/D66 pairing, and its
 * density), for meshes with one element layer per ply (elem_mat picks the ply). */
int sem_laminate_solid_layered(int lay, const double* alpha_deg, const double* rho, const double* q,
                               double* coef_out);

/* --- mesh / index maps --------------------------------------------------- */
/* Uploads connectivity and coordinates (main_flat_shell_force_excit_c.m:53-61,
 * solid engine :19-22), computes the GLL operators, per-element geometry
 * (force_excit_c.m:328-416; solid :480-604), the node->(element,point) inverse
 * map and the 12-basket colouring (nodesMaps_Fiborek.m:1-27) that fixes the
 * summation order of the assembly.  n_xy/n_z = GLL nodes per direction (n_z = 1
 * for shells). */
int sem_upload_mesh(sem_ctx* ctx, int kind, int n_xy, int n_z, int64_t n_el,
                    const int32_t* nodes, int64_t n_nodes, const double* coords);
/* A batch of independent cases as one mesh (no counterpart in the reference, whose
 * drivers loop over cases): case k has elements el_ptr[k]..el_ptr[k+1] of `nodes` (its own
 * 0-based node ids) and nodes node_ptr[k]..node_ptr[k+1] of `coords`.  The union numbers
 * case k's node v as node_ptr[k] + v, so DOFs, outputs and excitation address cases by
 * that offset.  Patches never mix cases; one run steps every case. */
int sem_upload_batch(sem_ctx* ctx, int kind, int n_xy, int n_z, int n_cases, const int64_t* el_ptr,
                     const int32_t* nodes, const int64_t* node_ptr, const double* coords);
/* Number of basket rows (ceil(n_el*n_loc/12)). */
int sem_basket_rows(const sem_ctx* ctx, int64_t* rows);
/* I_G / I_L of nodesMaps_Fiborek.m, 1-based, column-major [rows x 12], 0 = empty. */
int sem_get_baskets(const sem_ctx* ctx, int32_t* I_G, int32_t* I_L);
/* 1 if every element has a constant Jacobian (affine), else 0. */
int sem_mesh_is_affine(const sem_ctx* ctx, int* affine);

/* Element-kernel choice, the knob the autotuner turns (no counterpart in the reference):
 * variant 0 = point-wise GLL contractions (shells: 16-element patches, solids: 4),
 * variant 1 = line-wise contractions (shells: 16 or 8 elements per patch, default 16;
 *             solids: 4 or 8, default 8),
 * variant 2 = contractions as bf16 matrix-core products with FP32 accumulation and
 * hi/lo operand splitting (shells of order 4 and 5, FP32 form only),
 * variant 3 = line-wise contractions staged one direction at a time, gradients, stresses
 * and forces in registers (solids: 4 or 8 elements per patch, default 8).  patch_elems
 * (0 = the variant's default) takes effect at the next sem_upload_mesh; the variant from
 * the next run.  Results agree to rounding between variants. */
int sem_set_kernel(sem_ctx* ctx, int variant, int patch_elems);
int sem_get_kernel(const sem_ctx* ctx, int* variant, int* patch_elems);

/* --- material ------------------------------------------------------------ */
/* Per-material coefficient rows (SEM_SHELL5: 23 as sem_laminate_shell;
 * SEM_SOLID3: 14 as sem_laminate_solid_smeared), elem_mat[n_el] picks a row
 * per element (the reference fills its per-point arrays element by element,
 * composite_plate_c_delam.m:41-87, solid engine :69-95,492-530).  stiff_scale /
 * rho_scale (nullable = 1) scale each element's stiffness / mass coefficients.
 * Also assembles the lumped mass (force_excit_c.m:497-503; mass3D_v2.m:17-64). */
int sem_set_material(sem_ctx* ctx, int n_mat, int n_coef, const double* coef,
                     const int32_t* elem_mat, const double* stiff_scale,
                     const double* rho_scale);
/* Damping factor per field (dv, force_excit_c.m:517-525): cg = dv .* mg. */
int sem_set_damping(sem_ctx* ctx, const double* dv_field);
/* Lumped mass per DOF (the reference's mg before scaling by a0). */
int sem_get_mass(const sem_ctx* ctx, double* mass);

/* --- excitation, outputs, frames ----------------------------------------- */
/* F_ext = sum_k amp[k] e_{dofs[k]} * st(nn)  (Fi*st(nn), force_excit_c.m:582,687;
 * Fa*st(nn), multi_pzt_c.m:232-235,677).  st[nft] is indexed by the 1-based step. */
int sem_set_excitation(sem_ctx* ctx, int64_t n_src, const int64_t* dofs,
                       const double* amp, int64_t nft, const double* st);
/* Several signals: F_ext = sum_k amp[k] e_{dofs[k]} * st[sig[k]*nft + nn-1], with
 * st[n_sig x nft] row-major and sig nullable (= all 0): the solid engine's
 * four-actuator form F1*st1(nn)+F2*st2(nn)+F3*st3(nn)+F4*st4(nn)
 * (main_spec3Dcuda_global_cpu_gpu3_wave_sens_opt_single_PCT_mat.m:1034, commented out
 * there) for any number of actuators.  sem_set_excitation is this call with n_sig = 1;
 * sem_adjoint drives its reverse sweep with it. */
int sem_set_excitation_multi(sem_ctx* ctx, int64_t n_src, const int64_t* dofs,
                             const double* amp, const int32_t* sig, int n_sig,
                             int64_t nft, const double* st);
/* Output j = sum_{k in ptr[j]..ptr[j+1]} w[k] * q(dofs[k]) per step: point outputs
 * (displ(nn,:) = U(outputs), force_excit_c.m:701-705) have one entry of weight 1;
 * linear sensor functionals (PZT voltage) have many.  n_out = 0 removes the outputs. */
int sem_set_outputs(sem_ctx* ctx, int n_out, const int64_t* ptr, const int64_t* dofs,
                    const double* w, int quantity);
/* Frames at 1-based steps frame_nn[n_frames] (ascending), quantity SEM_Q_*
 * (force_excit_c.m:710-730 'velocity'/'displacement'/'acceleration'). */
int sem_set_frames(sem_ctx* ctx, int n_frames, const int64_t* frame_nn, int quantity);
/* Divergence guard evaluated at frame steps (force_excit_c.m:712-730): trips when
 * max over `field` (-1 = all fields) of U (post = 0: before the update, 1: after)
 * exceeds `limit`, or when U^{n+1}[2] is NaN.  limit <= 0 disables. */
int sem_set_guard(sem_ctx* ctx, double limit, int field, int post);

/* --- state ---------------------------------------------------------------- */
/* U^n and U^{n-1} (nullable = zeros); also rewinds frame/trace counters. */
int sem_set_state(sem_ctx* ctx, const double* u, const double* uold);
int sem_get_state(const sem_ctx* ctx, double* u, double* uold);

/* --- the hot path ----------------------------------------------------------- */
/* Time loop for 1-based steps nn = nn_first..nn_last (force_excit_c.m:592-862:
 * gather, element forces, assembly, central-difference update, sensors, frames);
 * 1 <= nn_first <= nn_last, else SEM_E_ARG.  A run continues from the context's state
 * (zero after an upload; sem_set_state sets it), so consecutive runs over consecutive
 * step ranges equal one run over their union.
 * traces: column-major [ld_traces x n_out], row nn-1 written for each step
 * (nullable); frames: [n_frames x dof] (nullable); t_frames[n_frames] (nullable).
 * *nn_done = last completed step.  Returns SEM_E_DIVERGED when the guard tripped
 * (rows after the tripping step are left zero, as the reference returns there).
 * *seconds (nullable) = wall time of the stepping loop (the reference's toc). */
int sem_run(sem_ctx* ctx, double dt, int64_t nn_first, int64_t nn_last, int precision,
            double* traces, int64_t ld_traces, double* frames, double* t_frames,
            int64_t* nn_done, double* seconds);
/* Per-kernel timing: runs steps nn_first..nn_first+nsteps-1 launched one by one with
 * HIP events around every kernel on the context's stream (no graph), and returns
 * the mean duration in ms of the element kernel, the node kernel and the post-step
 * kernel in ms_out[3].  Advances the state like sem_run; the reference's analogue is
 * its per-step tic/toc (main_flat_shell.m:652,918-924). */
int sem_profile(sem_ctx* ctx, double dt, int64_t nn_first, int64_t nsteps, int precision,
                double* ms_out);
/* Pure operator F_int = K u (force_excit_c.m:596-668), u and f of length dof. */
int sem_apply_K(sem_ctx* ctx, const double* u, double* f, int precision);
/* sem_apply_K for callers that cannot allocate device memory (code under graph capture,
 * tight serving loops, harnesses that forbid hipMalloc); no counterpart in the reference.
 * u and f are DEVICE pointers on the context's device, [dof] doubles in the caller's DOF
 * order; f receives the same values as sem_apply_K (0 on nodes of no element).  The call
 * allocates nothing: the internal-order copies and the boundary-node partial sums live in a
 * fixed device array of 128 MiB per device (held from sem_create until the process ends),
 * so meshes whose 2*dof + n_partials doubles exceed it are refused with SEM_E_ARG.  Work
 * runs on the context's stream and the call returns when f is written.  Every call on one
 * device shares that array: calls that overlap on one GPU (contexts on several host
 * threads) race and must be serialised by the caller. */
int sem_apply_K_noalloc(sem_ctx* ctx, const double* u, double* f, int precision);

/* GLL nodes and weights for n points on [-1, 1] (gll.m:1-39): the reference's literal table for
 * n = 2..10, bit for bit; the closed form beyond (n <= 64). */
int sem_gll(int n, double* x, double* w);

/* --- PZT transducers (set-up; the time loop sees sources and functionals) ---- */
/* Piezoelectric constants from the strain-charge form: S [6x6] compliance, d [3x6]
 * charge constants, epsT [3x3] (row-major, columns ex ey ez gxz gyz gxy).
 * variant 1 = solid3D/pzt_const.m:1-25 (g = epsT - e d^T, Q rearranged to
 * ex ey ez gxy gxz gyz); variant 0 = flat_shell_engine/pzt_const.m:1-37 (d rotated by
 * theta_deg, g = epsT, Q = S^-1 as given).  Outputs Q [6x6], e [3x6] (columns rearranged
 * to ex ey ez gxy gxz gyz), g [3x3]; e31_e32_g33 (nullable) = e(3,1), e(3,2) before the
 * rearrangement and g(3,3) (the flat shell's sensor constants). */
int sem_pzt_const(int variant, const double* S, const double* d, const double* epsT,
                  double theta_deg, double* Q, double* e, double* g, double* e31_e32_g33);
/* Actuator equivalent forces of a transducer meshed as n_pel hexahedra (pnodes
 * [n_pel x n_xy^2 n_z], bottom face first) over nodes with coords [n_nodes x 3]: element
 * matrices of coupl3D_v2.m (frame = 1) or coupl3D_global.m (frame = 0), V0 on the top
 * electrode, 0 on the bottom one, induced interior potential, fa = Kfiu^T phi
 * (main_spec3Dcuda_global_cpu_gpu3_wave_sens_opt_single_PCT_mat.m:198-293).
 * fa [3 n_nodes] (zero off the transducer). */
int sem_pzt_actuator(int n_xy, int n_z, int64_t n_pel, const int32_t* pnodes, int64_t n_nodes,
                     const double* coords, const double* e, const double* g, int frame,
                     double V0, double* fa);
/* Open-circuit sensor voltage as a functional of the displacements, voltage = w . U
 * (same file :304-388 and :1056-1067: fio = -KfifiO \ (KfiuO U), mean of the last
 * n_xy^2 potentials).  w [3 n_nodes]. */
int sem_pzt_sensor(int n_xy, int n_z, int64_t n_pel, const int32_t* pnodes, int64_t n_nodes,
                   const double* coords, const double* e, const double* g, int frame, double* w);
/* Flat-shell actuator (main_flat_shell_multi_pzt_c.m:158-235): a 3D transducer of n_z GLL
 * levels and thickness t_pzt over the actuator's shell elements act_nodes [n_act x n_xy^2]
 * (connect_pzt3D_to_plate_fun.m), coupl3D_global matrices, induced potential; in-plane
 * forces of the footprint's boundary nodes (find_boundary_nodes.m) plus moments
 * (h/2 + t_pzt/2) F.  fa [5 n_nodes]. */
int sem_pzt_shell_actuator(int n_xy, int n_z, int64_t n_act, const int32_t* act_nodes,
                           int64_t n_nodes, const double* coords, double t_pzt, double h,
                           const double* e, const double* g, double V0, double* fa);
/* Shell coefficient row (23, sem_laminate_shell order) with the transducer layer added
 * (composite_plate_c_delam_pzt.m:175-215); Q as returned by sem_pzt_const variant 0. */
int sem_laminate_shell_add_pzt(const double* row, double h, double t_pzt, double rho_pzt,
                               const double* Q, double* out);
/* Flat-shell sensor voltage as a functional of the state, voltage = w . U, over the
 * points of the sensor's elements elems[n_sel] (caller's element ids) of the uploaded
 * mesh: mean of t_pzt/g33 (e31 (u,x + h/2 phix,x) + e32 (v,y + h/2 phiy,y)) with the
 * reference's invJ pairing (main_flat_shell_multi_pzt_c.m:618-623).  w [dof]. */
int sem_shell_pzt_sensor(const sem_ctx* ctx, int64_t n_sel, const int32_t* elems, double t_pzt,
                         double h, double e31, double e32, double g33, double* w);

/* --- adjoint (no counterpart in the reference, which only integrates forward) -- */
/* Discrete adjoint of sem_run over steps nn_first..nn_last from the current state:
 * given dJ/dtraces as dJ [ld x n_out] column-major (row nn-1 per step, the layout of
 * sem_run's traces, ld >= nn_last), returns dJ/dstiff_scale and dJ/drho_scale per
 * element in g_stiff[n_el] / g_rho[n_el] (nullable), at the scales set by
 * sem_set_material.  The forward states are checkpointed every seg_len steps
 * (<= 0: ceil(sqrt(steps))) and recomputed one segment at a time for the reverse
 * sweep, which runs the same element kernels (K is symmetric).  Outputs, frames and the
 * guard are not evaluated; the state is left at step nn_last as after sem_run.
 * *seconds (nullable) = wall time of both sweeps. */
int sem_adjoint(sem_ctx* ctx, double dt, int64_t nn_first, int64_t nn_last, int precision,
                const double* dJ, int64_t ld, double* g_stiff, double* g_rho,
                int64_t seg_len, double* seconds);

#ifdef __cplusplus
}
#endif
#endif /* SEM_HIP_H */
