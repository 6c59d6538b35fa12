/* C ABI of libanneal_hip.so: the MI355X (gfx950) SIMP hot path of Anneal.
 *
 * The reference (ImpurestTadpole/anneal) is pure Python with no FFI; these
 * entry points are what its hot-path calls bind to.  Each declaration names
 * the reference code it replaces.
 *
 * Conventions
 *  - Every function returns 0 on success and a non-zero status on failure;
 *    anneal_last_error() then returns a message for the calling thread.
 *  - Pointers named *_dev are device pointers (HIP); *_host are host pointers.
 *  - Voxel grid nx*ny*nz, element e = i + nx*(j + ny*k); node
 *    n = i + (nx+1)*(j + (ny+1)*k) (anneal/fem.py:87-127).
 *  - Nodal vectors are SoA: component c of right-hand side r starts at
 *    (3*r + c) * ldn, where ldn (>= nnode) is reported by anneal_fem_info.
 *    The reference's interleaved dof 3n+c (fem.py:143) maps to [c*ldn + n].
 *  - Element fields cover the full grid; flags bit0 = active, bit1 = frozen.
 *  - stream is a hipStream_t (may be NULL for the default stream); all work of
 *    a context is ordered on it.  Contexts are not thread-safe; use one per
 *    thread (the reference runs one job per worker thread, gui/server.py:52-56).
 */
#ifndef ANNEAL_HIP_H
#define ANNEAL_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Message of the last failed call on this thread ("" if none). */
const char* anneal_last_error(void);
/* ABI version (increments on incompatible changes). */
int anneal_abi_version(void);

/* ---- hex8 voxel solve: replaces anneal/fem.py FEM ------------------------ */

/* hex8_matrices(nu, h)[0] (fem.py:71-84) as the library computes it (host
 * only, no device needed): ke_host receives the 24x24 stiffness for E = 1 in
 * the library's corner order q = a + 2b + 4c (corner at (a,b,c)), dof 3q+comp;
 * coef_host receives the 8 Walsh-basis coefficients the kernels use
 * (normal A-B, B, shear, pair C, pair D, triple a-b, b, trilinear), already
 * divided by 64.  Fails if ke does not have the block pattern they assume. */
int anneal_ke_host(double nu, double h, double* ke_host, double* coef_host);

/* FEM.__init__ (fem.py:133-196): active voxels and fixed dofs define the free
 * set.  active_host: nelem bytes (non-zero = active).  fixed_host: nnode bytes,
 * bit c set when dof 3n+c is fixed.  nrhs_max: largest right-hand-side block
 * (load cases) solved at once.  Builds the multigrid hierarchy, coarsening
 * until a level has at most max_coarse_nodes nodes (0 = the library default;
 * that level is solved with a dense FP64 inverse, so keep it <= 4096). */
int anneal_fem_create(void** ctx, int nx, int ny, int nz, double h, double nu, const uint8_t* active_host,
                      const uint8_t* fixed_host, int nrhs_max, int max_coarse_nodes, void* stream);
int anneal_fem_destroy(void* ctx);
/* Leading dimension of nodal vectors, number of multigrid levels, and the
 * free-dof mask (nnode bytes, bit c = dof 3n+c free) copied to mask_host if
 * not NULL. */
int anneal_fem_info(void* ctx, int64_t* ldn, int* levels, uint8_t* mask_host);
/* Multigrid level l (0 = finest): grid dimensions and FP32 apply tile shape id. */
int anneal_fem_level_info(void* ctx, int level, int* nx, int* ny, int* nz, int* tile);
/* FEM.assemble(Ee) (fem.py:198-201): Ee_dev holds nelem FP64 stiffness
 * scales (0 on inactive voxels).  No matrix is formed: this sets the operator
 * and rebuilds the multigrid levels. */
int anneal_fem_set_Ee(void* ctx, const double* Ee_dev);
/* Refresh the expensive part of the multigrid hierarchy (the coarsest
 * level's dense inverse) for the current Ee: the counterpart of the
 * reference's AMG rebuild (fem.py:207-216).  Between rebuilds the solve keeps
 * the previous coarse inverse (still a valid SPD coarse solve), unless the
 * option "rebuild_each_solve" is 1. */
int anneal_fem_rebuild(void* ctx);
/* y = K u on free dofs (fixed/unused dofs give 0): the product the reference
 * forms with its CSR matrix (fem.py:201), FP32 and FP64 variants. */
int anneal_fem_apply_f32(void* ctx, const float* u_dev, float* y_dev, int nrhs);
int anneal_fem_apply_f64(void* ctx, const double* u_dev, double* y_dev, int nrhs);
/* FEM.solve (fem.py:206-232): K U = F for nrhs right-hand sides at once, stop
 * when ||F - K U||_2 <= rtol ||F||_2 (FP64 true residual) for every one.
 * U_dev holds the initial guess when use_x0.  iters_host[r] receives the PCG
 * iterations of rhs r.  Returns 0, or 3 = not converged within maxiter,
 * 4 = non-finite result (the reference raises SolverError for both). */
int anneal_fem_solve(void* ctx, const double* F_dev, double* U_dev, int nrhs, double rtol, int maxiter, int use_x0,
                     int* iters_host);
/* Tile-shape autotuner of the FP32 apply (no reference counterpart: the
 * reference has no kernels).  A context on a grid of at least 4096 elements
 * picks its shapes when it is created: those the process already chose for
 * the same device and multigrid hierarchy, else by this timing sweep.
 * variant >= 0 selects a shape; variant == -2
 * only reports the shape in use; any other negative value times all shapes
 * on this grid and keeps the fastest (for later contexts too).  *chosen_host receives the
 * shape in use; us_host (may be NULL; room for 16 floats) receives *n_host per-shape times in
 * microseconds, where *n_host is the number of shapes. */
int anneal_fem_autotune(void* ctx, int variant, int* chosen_host, float* us_host, int* n_host);
/* Pin the FP32 apply tile shape (an id as reported by anneal_fem_level_info)
 * of multigrid level `level` (no reference counterpart): reproduces a tile
 * choice of the autotuner, e.g. under a profiler whose per-dispatch overhead
 * hides the differences the autotuner measures. */
int anneal_fem_set_level_tile(void* ctx, int level, int variant);
/* Solver options (no reference counterpart; the reference's solver has none):
 * "graphs" (1 = replay the V-cycle and each PCG iteration from captured
 * hipGraphs, 0 = launch the same kernels one by one; default 0, which
 * measured 1-2 % faster), "smoother_degree" (Chebyshev degree per side on the fine level,
 * >= 1; default 3), "smoother_degree_coarse" (levels >= 1; default 3),
 * "smoother_degree_deep" (levels >= 2; 0 = as the coarse levels),
 * "deep_escalate_iters" (a solve averaging more PCG iterations per rhs than
 * this, while smoother_degree_deep is 0, sets the deep levels to
 * "deep_escalate_degree" for the rest of the context; defaults 20 and 13,
 * 0 = off), "cheb_ratio"
 * (lambda_max / lambda_min of the smoothed interval; default 15),
 * "coarsen_power" (q of the power mean
 * that coarsens the stiffness scales, (0, 1]), "fuse_first_step" (1 = the first
 * two smoothing steps in one kernel on every level, 0 = never),
 * "fuse_first_level" (fuse from this level down), "inner_tol_factor" (the
 * FP32 inner PCG stops at factor * rtol * |F| / |r|; default 0.85), "inner_tol_floor"
 * (largest relative residual reduction asked of one inner PCG), "inner_max"
 * (iterations per inner PCG), "rebuild_each_solve" (1 = every solve after a
 * new Ee refreshes the coarse inverse, 0 = only anneal_fem_rebuild does).
 * Unknown names fail. */
int anneal_fem_set_option(void* ctx, const char* name, double value);
/* One V-cycle of the multigrid preconditioner on FP32 vectors (z = M r). */
int anneal_fem_vcycle(void* ctx, const float* r_dev, float* z_dev, int nrhs);
/* One fine-level Chebyshev smoothing step of that V-cycle on FP32 vectors, in
 * place: y <- M (u + c1 (u - y) + c2 D^-1 (b - K u)) (the post-smoothing
 * kernel, the largest share of the solve's kernel time; no reference
 * counterpart: for measurement). */
int anneal_fem_smooth_step_f32(void* ctx, const float* u_dev, const float* b_dev, float* y_dev, int nrhs);
/* FEM.element_energy (fem.py:239-242): ce[r*nelem + e] = u_e^T ke u_e, E = 1. */
int anneal_fem_element_energy(void* ctx, const double* U_dev, int nrhs, double* ce_dev);
/* Sensitivities of optimize.py:118,125-142 fused: with
 * dce_r = -p xPhys^(p-1) (1-emin) E0 phi ce_r, dc = min(sum_r coef_r dce_r, 0),
 * writes dcdx = dc*dxp and dvdx = phi*dxp (0 on inactive voxels).
 * coef_dev: nrhs FP64 weights (w for "sum", the p-norm chain factor for "worst"). */
int anneal_fem_sensitivity(void* ctx, const double* U_dev, int nrhs, const uint8_t* flags_dev, const double* xphys_dev,
                           const double* dxp_dev, const double* phi_dev, const double* coef_dev, double p, double emin,
                           double E0, double* dcdx_dev, double* dvdx_dev);

/* ---- density filter and OC update: replaces anneal/optimize.py ---------- */

/* Filter(grid, active, r_vox) (optimize.py:24-34); flags_host: nelem bytes. */
int anneal_filter_create(void** ctx, int nx, int ny, int nz, double r_vox, const uint8_t* flags_host, void* stream);
int anneal_filter_destroy(void* ctx);
/* Filter.forward / Filter.backward (optimize.py:41-47) on full-grid FP64
 * arrays; inactive voxels read as and are written as 0. */
int anneal_filter_forward(void* ctx, const double* x_dev, double* y_dev);
int anneal_filter_backward(void* ctx, const double* d_dev, double* y_dev);
/* The two backward passes of the sensitivities (optimize.py:142-143: dc and
 * dv) in one launch: y1 = backward(d1), y2 = backward(d2), the same bits as two
 * anneal_filter_backward calls.  stats_dev (may be NULL) receives the design
 * statistics the OC bracket is formed from, in the layout of the statistics
 * array: [3] the number of design voxels, [4] the sum of -y1, [5] the sum of y2
 * over them, [0..2] = 0. */
int anneal_filter_backward_pair(void* ctx, const double* d1_dev, double* y1_dev, const double* d2_dev,
                                double* y2_dev, double* stats_dev);
/* physical(x, with_grad=True) (optimize.py:81-93): filter, Heaviside
 * projection with sharpness beta (identity for beta <= 0), frozen -> 1.
 * vol_dev (may be NULL) receives sum(phi * xPhys). */
int anneal_filter_physical(void* ctx, const double* x_dev, double beta, const double* phi_dev, double* xphys_dev,
                           double* dxp_dev, double* vol_dev);
/* The OC bisection loop (optimize.py:144-161) on the device, starting from the
 * bracket [lo, hi].  The trial multipliers, the decisions vol > vtarget and
 * the stop hi/lo < 1+1e-9 are the reference's; since vol(lam) is
 * non-increasing, trials outside a certified bracket of its root are decided
 * without a volume evaluation.  The multipliers evaluated follow a safeguarded
 * secant in log(lam): the first at hint (0 = none: the reference's first
 * trial), the second a Newton step with the slope d(vol)/d(log lam) that the
 * context's previous bisection measured, or, without one, a log step of
 * hint_width (<= 0: 0.005) toward the root.  These change which multipliers
 * are evaluated, never the trial decisions or the result.  xnew_dev receives
 * the last trial design, *steps_host the
 * number of trials, *evals_host the volume evaluations spent (may be NULL) and
 * *lam_last_host the last trial multiplier (may be NULL). */
int anneal_oc_bisect(void* ctx, const double* x_dev, const double* dcx_dev, const double* dvx_dev,
                     const double* phi_dev, double beta, double move, double lo, double hi, double vtarget,
                     double hint, double hint_width, double* xnew_dev, int* steps_host, int* evals_host,
                     double* lam_last_host);

/* anneal_oc_bisect in three calls (optimize.py:144-161), so that one host
 * thread can keep the bisections of several filter contexts in flight:
 * anneal_oc_begin queues the first volume evaluations and a read-back of the
 * bisection state on the context's stream and returns; anneal_oc_poll waits for
 * that read-back only and sets *done_host = 1 once the bisection has stopped,
 * else queues the next evaluations and sets 0; once done, anneal_oc_end
 * queues the final design into xnew_dev and returns the counts as
 * anneal_oc_bisect does.  Launches and results equal anneal_oc_bisect's; one
 * bisection per context at a time. */
int anneal_oc_begin(void* ctx, const double* x_dev, const double* dcx_dev, const double* dvx_dev,
                    const double* phi_dev, double beta, double move, double lo, double hi, double vtarget, double hint,
                    double hint_width, double* xnew_dev);
/* anneal_oc_begin with the bracket formed on the device from the design
 * statistics that anneal_design_stats writes, where stats_dev[3] is the design
 * count, [4] the sum of -dcx and [5] the sum of dvx over design voxels: scale = max(mean(-dcx) /
 * max(mean(dvx), 1e-30), 1e-300) (1 without design voxels), lo = 1e-8 scale,
 * hi = 1e8 scale (optimize.py:146-148), and the first evaluation at
 * hint_ratio * scale (<= 0: none), so the caller needs no wait for the
 * statistics before the bisection starts.  Continue with anneal_oc_poll and
 * anneal_oc_end. */
int anneal_oc_begin_stats(void* ctx, const double* x_dev, const double* dcx_dev, const double* dvx_dev,
                          const double* phi_dev, double beta, double move, const double* stats_dev,
                          double hint_ratio, double vtarget, double hint_width, double* xnew_dev);
int anneal_oc_poll(void* ctx, int* done_host);
int anneal_oc_end(void* ctx, int* steps_host, int* evals_host, double* lam_last_host);

/* ---- element-wise pieces of simp() ---------------------------------------- */

/* Ee = E0 phi (emin + xPhys^p (1-emin)) on active voxels, else 0 (optimize.py:118). */
int anneal_stiffness_law(const double* xphys_dev, const double* phi_dev, const uint8_t* flags_dev, double E0,
                         double emin, double p, int64_t n, double* Ee_dev, void* stream);
/* Reductions used by simp() for its history (optimize.py:145, 162-167):
 * out_dev[0] = sum phi*xPhys over active, [1] = #grey design voxels,
 * [2] = max |xnew - xold| over design, [3] = #design voxels,
 * [4] = sum(-dcx) over design, [5] = sum(dvx) over design.  Any input may be
 * NULL (its outputs are then 0).  work_dev: 3072 doubles. */
int anneal_design_stats(const double* xold_dev, const double* xnew_dev, const double* xphys_dev,
                        const double* phi_dev, const double* dcx_dev, const double* dvx_dev,
                        const uint8_t* flags_dev, int64_t n, double* out_dev, double* work_dev, void* stream);
/* FP64 dot products of nodal vectors: out_dev[r] = sum over the 3 components
 * of a_r . b_r (compliance f.u, optimize.py:125).  work_dev: 512*nrhs doubles. */
int anneal_dot_f64(const double* a_dev, const double* b_dev, int64_t nnode, int64_t ldn, int nrhs, double* out_dev,
                   double* work_dev, void* stream);
/* anneal_dot_f64 without the work_dev argument: no counterpart in the
 * reference.  For callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc): the
 * partial sums live in a static 256 KiB device array of the library, so this
 * call allocates nothing.  nrhs <= 64 (larger fails with status 1).  Calls
 * that overlap on one device share that array and race: order them on one
 * stream.  Same kernels and results as anneal_dot_f64. */
int anneal_dot_f64_noalloc(const double* a_dev, const double* b_dev, int64_t nnode, int64_t ldn, int nrhs,
                           double* out_dev, void* stream);

/* FEM.stress / FEM.von_mises (fem.py:244-263) for nodal solutions U_dev
 * (FP64 [nrhs][3][ldn]): sig_dev[r][e][6] = E_e D B0 u_e at the element centre
 * (Voigt xx yy zz xy yz zx), vm_dev[r][e] (may be null) its von Mises stress.
 * B0_host (6x24, dof 3q+c in the corner order local_host[8][3]) and D_host
 * (6x6) are hex8_matrices' outputs (fem.py:71-84); E_dev is a per-element
 * modulus over the full grid or null for the scalar E.  Inactive voxels get 0. */
int anneal_fem_stress(void* ctx, const double* U_dev, int nrhs, const double* B0_host, const double* D_host,
                      const int* local_host, const double* E_dev, double E, double* sig_dev, double* vm_dev);

/* ---- result fields on a display mesh (anneal/fields.py) ------------------ */

/* fields.corner_values (fields.py:17-37) after the npz has been read
 * (fields.load, fields.py:10-14): out_dev[3*nface][4] float32 =
 * [dx, dy, dz, vonMises] per triangle corner, in triangle order.
 *   u_dev   [nnode][3] float32 nodal displacement ("u::case", node index
 *           i + (nx+1)(j + (ny+1)k));
 *   vm_dev  [nelem] float32 element von Mises ("vm::case");
 *   origin_host  3 doubles ("origin"); h ("h"); nx, ny, nz ("shape");
 *   tri_dev [nface][3][3] float64 corner points (mesh.triangles);
 *   nrm_dev [nface][3] float64 face normals (mesh.face_normals).
 * Displacement: trilinear, coordinates clamped to the node box; stress: 2x2x2
 * grey dilation of vm, order-0 lookup half a voxel inside the surface. */
int anneal_fields_corner_values(const float* u_dev, const float* vm_dev, int nx, int ny, int nz, double h,
                                const double* origin_host, const double* tri_dev, const double* nrm_dev,
                                int64_t nface, float* out_dev, void* stream);

/* ---- learned preconditioner: 3-D U-Net (no counterpart in the reference) ---
 * Channels-last bf16 tensors on node grids (voxel i + nx (j + ny k), sample
 * stride nx ny nz); convolutions are BF16 MFMA implicit GEMMs with FP32
 * accumulation.  Flat parameter order: for the encoder levels (conv a, conv b),
 * then the decoder levels from the finest (conv a, conv b), then the 1x1 output
 * convolution; each layer W[co][tap][ci] (tap = dx+1 + 3(dy+1) + 9(dz+1))
 * followed by b[co].  Used by anneal_amd/learned.py. */
int anneal_unet_create(void** ctx, int cin, int cout, const int* widths, int levels, void* stream);
int anneal_unet_destroy(void* ctx);
int anneal_unet_nparams(void* ctx, int64_t* n);
int anneal_unet_get_params(void* ctx, float* host);
int anneal_unet_set_params(void* ctx, const float* host);
int anneal_unet_get_grads(void* ctx, float* host);
/* out_dev[B][V][cout] FP32 = net(in_dev[B][V][cin] bf16), V = nx ny nz nodes. */
int anneal_unet_forward(void* ctx, const uint16_t* in_dev, int B, int nx, int ny, int nz, float* out_dev);
/* Accumulate parameter gradients of the last forward for dL/dout (FP32). */
int anneal_unet_backward(void* ctx, const float* dout_dev);
int anneal_unet_zero_grad(void* ctx);
int anneal_unet_adam(void* ctx, float lr, float beta1, float beta2, float eps, int step, float weight_decay);
/* loss_dev = mean_b sum (o - t)^2 / sum t^2 over each sample; dout_dev (may be
 * null) its gradient.  work_dev: 130 B doubles. */
int anneal_unet_rel_mse(const float* o_dev, const float* t_dev, int B, int nx, int ny, int nz, int c, float* dout_dev,
                        double* loss_dev, double* work_dev, void* stream);
/* Node-aligned grid transfer of bf16 channels-last tensors (c a multiple of 8):
 * restrict y[I] = s sum_d w(d) x[2I+d]; prolong y[i] (+)= s sum_parents w x. */
int anneal_unet_restrict(const uint16_t* x_dev, int B, int fx, int fy, int fz, int c, uint16_t* y_dev, float s,
                         void* stream);
int anneal_unet_prolong(const uint16_t* x_dev, int B, int cx, int cy, int cz, int c, uint16_t* y_dev, int fx, int fy,
                        int fz, float s, int accumulate, void* stream);

/* Attach a U-Net (anneal_unet_create with 8 input and 3 output channels) as a
 * stage of the solve (no counterpart in the reference).  With
 * z1 = gamma P N(R r), gamma = <r, z1> / <z1, K z1> (energy-optimal):
 * mode 0 off; 1 warm start (each inner PCG starts from x = z1); 2 two-stage
 * preconditioner z = z1 + V(r - K z1) in every iteration, with flexible CG.
 * The network runs on the node grid of multigrid level `level` (>= 1).  The network context must outlive its
 * use by the FEM context.  e_ref scales the stiffness input and the output. */
int anneal_fem_set_learned(void* fem, void* unet, int level, double e_ref, int mode);
/* z_dev = P^level N(R^level r_dev) on FP32 [nrhs][3][ldn] buffers (diagnostics). */
int anneal_fem_learned_stage(void* fem, const float* r_dev, float* z_dev, int nrhs);
/* Training pair for FP64 loads F_dev and solutions U_dev ([nrhs][3][ldn]) at
 * the current Ee: in_dev[nrhs][V][8] bf16 network input for r = F and
 * tgt_dev[nrhs][V][3] FP32 the solution injected onto the level, in output
 * units (V = nodes of that level). */
int anneal_fem_learned_pair(void* fem, const double* F_dev, const double* U_dev, int nrhs, int level, double e_ref,
                            uint16_t* in_dev, float* tgt_dev);

/* ---- several independent problems solved together (no counterpart in the
 * reference, whose FEM.solve handles one grid at a time, fem.py:206-232) ----
 * A batch groups FEM contexts of any grid sizes that share one stream.  One
 * PCG iteration of a problem runs on the problem's own stream (replayed from a
 * hipGraph recorded once per host slot of its norms when the member's "graphs"
 * option is on), forked from the shared
 * stream once per refinement pass and joined back at its end; each problem's
 * next iteration is launched before the host waits for its current one, and
 * the host reads the norms of every problem's iteration in one round.
 * F[p], U[p]: FP64 [nrhs[p]][3][ldn_p];
 * status[p] as anneal_fem_solve (0, 3 not converged, 4 non-finite);
 * iters[p * 64 + r] the PCG iterations of rhs r.  Plain multigrid only. */
int anneal_batch_create(void** batch, void* const* fems, int nprob);
/* A problem group: nmember FEM contexts of one grid shape (same h, Poisson
 * ratio and multigrid levels) solved as ONE block of right-hand sides, problem
 * p owning rhs [p k, (p + 1) k) of an anneal_fem_solve with nrhs = k * nmember.
 * Every kernel of the solve (apply, smoothers, transfers, coarse solve, PCG
 * vectors) covers all problems in one launch, reading each problem's
 * operator from its member context.  The members keep their own stiffness
 * (anneal_fem_set_Ee) and coarse rebuilds, use the group's stream and must
 * outlive it; destroy the group with anneal_fem_destroy.  A group handle can
 * itself be a member of anneal_batch_create. */
int anneal_fem_group_create(void** ctx, void* const* members, int nmember, int nrhs_max, void* stream);
int anneal_batch_destroy(void* batch);
int anneal_batch_solve(void* batch, const double* const* F_dev, double* const* U_dev, const int* nrhs, double rtol,
                       int maxiter, int use_x0, int* status, int* iters);

/* ---- voxel passes before and after the SIMP loop (anneal/problem.py,
 * anneal/export.py).  Element fields cover the full nx*ny*nz grid (voxel
 * i + nx (j + ny k)); node fields (nx+1)(ny+1)(nz+1).  Masks are one byte per
 * voxel (0/1).  work_dev arguments are 4 device int64 of scratch.  Entry
 * points that return counts synchronise the stream. ----------------------- */

/* Anchor regions as signed-distance primitives (anneal/kernel.py:40-125):
 * kind_dev[nprim] (0 axis box: c, half; 1 framed box: o, x, y, z axes,
 * z0 + size_z / 2, half; 2 sphere: c, r; 3 capped cylinder or pipe: a, unit
 * axis, L, r, r_in), prm_dev[nprim][16], anchor a = primitives
 * [first_dev[a], first_dev[a+1]) combined by minimum (Union, kernel.py:118-121),
 * selected where that minimum is <= thr_dev[a].  at_nodes = 1: the points are
 * the nodes of voxels with emask set at origin + h (i, j, k) (Anchor.nodes on
 * the candidate nodes of Problem.boundary / _connected, problem.py:103-113,
 * 276-300); at_nodes = 0: voxel centres origin + h (ijk + 0.5) with emask set
 * (the anchor groups of export.keep_load_paths, export.py:41-45).
 * bits_dev[point] bit a = selected; counts_dev[32] int64 per-anchor totals. */
int anneal_region_select(int nx, int ny, int nz, double h, const double* origin_host, const int* kind_dev,
                         const double* prm_dev, const int* first_dev, const double* thr_dev, int nanchor,
                         int at_nodes, const uint8_t* emask_dev, uint32_t* bits_dev, long long* counts_dev,
                         void* stream);
/* out_dev[a][p] = the anchor's signed distance at pts_dev[p][3] (Shape.sdf). */
int anneal_region_sdf(int64_t npts, const double* pts_dev, const int* kind_dev, const double* prm_dev,
                      const int* first_dev, int nanchor, double* out_dev, void* stream);
/* Problem.boundary after the anchor selection (problem.py:302-326):
 * fixed_bits_dev[nnode] = OR of dofmask_dev[a] (bit d = dof d) over the anchors
 * selecting the node; F_dev[ncase][3 nnode] (the reference's interleaved dof
 * layout) = sum over anchors in order of load_dev[c][a][d] = v_d / count_a. */
int anneal_boundary_fields(int nx, int ny, int nz, const uint32_t* node_bits_dev, int nanchor,
                           const uint8_t* dofmask_dev, const double* load_dev, int ncase, uint8_t* fixed_bits_dev,
                           double* F_dev, void* stream);
/* ebits_dev[e] = OR of the bits of voxel e's 8 nodes where emask_dev[e]. */
int anneal_gather_node_bits(int nx, int ny, int nz, const uint32_t* node_bits_dev, const uint8_t* emask_dev,
                            uint32_t* ebits_dev, void* stream);
/* Face-connected components (scipy.ndimage.label with its default structure,
 * problem.py:278, export.py:49): label_dev[e] = smallest voxel index of the
 * component (-1 outside the mask); *ncomp_out = number of components. */
int anneal_voxel_components(int nx, int ny, int nz, const uint8_t* mask_dev, int32_t* label_dev,
                            long long* work_dev, int64_t* ncomp_out, void* stream);
/* Components holding a voxel with ebits & seed ("touch a fixed anchor",
 * problem.py:283-297; fixed_labels, export.py:51): *nkept_out of them;
 * out_mask_dev (may be null) = mask on those components; *linked_out = OR of
 * ebits over their voxels.  keep_dev: nelem bytes of scratch. */
int anneal_voxel_link(int nx, int ny, int nz, const uint8_t* mask_dev, const int32_t* label_dev,
                      const uint32_t* ebits_dev, uint32_t seed, uint8_t* keep_dev, uint8_t* out_mask_dev,
                      long long* work_dev, int64_t* nkept_out, uint32_t* linked_out, void* stream);
/* mask_dev[e] = f_dev[e] > t;  *count_out = number of e with f_dev[e] > t. */
int anneal_voxel_threshold(int64_t n, const double* f_dev, double t, uint8_t* mask_dev, void* stream);
int anneal_voxel_count_above(int64_t n, const double* f_dev, double t, long long* work_dev, int64_t* count_out,
                             void* stream);
/* export.output_field (export.py:8-14). */
int anneal_output_field(int64_t n, const double* x_dev, const double* fill_dev, const uint8_t* active_dev,
                        const double* solid_fill_dev, const double* void_fill_dev, const uint8_t* frozen_dev,
                        double* f_dev, void* stream);
/* export.choose_level (export.py:17-28): *level_out, *target_out = sum x fill.
 * work_dev: anneal_choose_level_work_bytes() bytes. */
int64_t anneal_choose_level_work_bytes(void);
int anneal_choose_level(int64_t n, const double* f_dev, const double* x_dev, const double* fill_dev, void* work_dev,
                        double* level_out, double* target_out, void* stream);
/* Shortest paths of export.keep_load_paths' repair stage (export.py:70-83,
 * scipy dijkstra min_only from the fixed groups): dist_dev[nelem] FP64 (inf
 * where unreachable or inactive), pred_dev[nelem] (-1 for sources / none).
 * stamp_dev: nelem int32, flag_dev: one int of scratch. */
int anneal_paths_sssp(int nx, int ny, int nz, const uint8_t* active_dev, const double* f_dev, double need,
                      const uint32_t* ebits_dev, uint32_t source_groups, double* dist_dev, int32_t* pred_dev,
                      int32_t* stamp_dev, int* flag_dev, int* sweeps_out, void* stream);
/* Start voxel of group `group` (export.py:86-87) and its path back to a source
 * (export.py:88-95) into path_dev (nelem int32); *len_out = 0 if unreachable. */
int anneal_paths_trace(int nx, int ny, int nz, const uint8_t* active_dev, const uint32_t* ebits_dev, int group,
                       const double* dist_dev, const int32_t* pred_dev, int32_t* path_dev, long long* work_dev,
                       int64_t* start_out, int64_t* len_out, double* dstart_out, void* stream);
/* Material added along a path (export.py:96-101). */
int anneal_paths_raise(int nx, int ny, int nz, const int32_t* path_dev, int64_t len, const uint8_t* active_dev,
                       const double* fill_dev, double need, double* f_dev, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* ANNEAL_HIP_H */
