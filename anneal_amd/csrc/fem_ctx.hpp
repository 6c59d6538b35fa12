// Device-resident hex8 voxel operator, geometric multigrid and mixed-precision PCG.
//
// Replaces FEM.__init__/assemble/solve of anneal/fem.py:133-232: instead of an
// assembled CSR and pyamg, the operator is (ke, Ee, dof mask) and the solve is
// PCG preconditioned by a rediscretised geometric V-cycle, with FP32 vectors,
// FP64 dot products and FP64 true-residual refinement to the requested rtol.
#pragma once
#include <map>
#include <tuple>
#include <vector>
#include "stencil.hpp"

namespace anneal {

class UNet;

// Coarsening stops once a level has at most this many nodes (dense FP64 inverse there).
constexpr int kDefaultCoarseNodes = 700;

// 24x24 hex8 stiffness for E = 1, cube of side h, 2x2x2 Gauss quadrature.
// Corner q = a + 2b + 4c for the corner at (a, b, c) in {0,1}^3, dof 3q + comp.
void hex8_stiffness(double nu, double h, double* ke /*576*/);
// Walsh-domain coefficients of ke; throws if the block pattern does not hold.
KeCoef<double> walsh_coefficients(const double* ke);
// lambda_max(ke) / ke_00: upper bound of lambda_max(D^-1 K) for any Ee >= 0.
double element_jacobi_bound(const double* ke);

struct Level {
  GridDims g{};
  double h = 1.0;
  uint8_t* mask = nullptr;  // device, per node
  float* Ee = nullptr;      // device, per element
  float* dinv = nullptr;    // device, per node
  float *x = nullptr, *x2 = nullptr, *b = nullptr, *r = nullptr;  // nrhs_max x 3 x ldn
  KeCoef<float> k{};
  double kdiag = 0.0;
  int tile = 0;  // FP32 apply tile shape for this level (autotuned)
};

struct SolveStats {
  int iters[64];
  int outer;
};

class FemContext {
 public:
  FemContext(int nx, int ny, int nz, double h, double nu, const uint8_t* active_host, const uint8_t* fixed_host,
             int nrhs_max, hipStream_t stream, int max_coarse_nodes);
  // A group: several problems of one grid shape (same h, Poisson ratio and
  // multigrid levels) solved as one block of right-hand sides, problem p owning
  // rhs [p * rpo, (p + 1) * rpo) with rpo = nrhs / members.  Every kernel of the
  // solve covers all problems in one launch, reading each problem's operator
  // (stiffness scales, masks, Jacobi diagonals, coarse inverse) from the member
  // contexts through per-level device tables.  The members keep their own
  // stiffness (set_Ee) and coarse rebuilds, share the group's stream and must
  // outlive it.
  FemContext(const std::vector<FemContext*>& members, int nrhs_max, hipStream_t stream);
  bool is_group() const { return !members_.empty(); }
  int group_size() const { return (int)members_.size(); }
  // group: rhs per problem for the next kernels (checks nrhs is a multiple of the group size)
  void set_rhs_count(int nrhs);
  ~FemContext();
  FemContext(const FemContext&) = delete;
  FemContext& operator=(const FemContext&) = delete;

  void set_Ee(const double* Ee_dev);
  // Refresh the coarsest level's dense inverse for the current Ee (the
  // expensive part of the hierarchy; the rediscretised levels and Jacobi
  // diagonals follow every set_Ee).  A stale inverse is still a symmetric
  // positive definite coarse solve, so PCG stays correct with it; only the
  // iteration count can grow.  This mirrors the reference's "rebuild" of its
  // AMG hierarchy (fem.py:206-216, optimize.py:120-124).
  void rebuild_coarse();
  bool rebuild_each_solve = true;  // solve() refreshes a stale coarse inverse
  void apply_f32(const float* u, float* y, int nrhs);
  void apply_f64(const double* u, double* y, int nrhs);
  // One fine-level Chebyshev smoothing step in place, y <- M (u + c1 (u - y) + c2 D^-1 (b - K u)), with
  // the coefficients of the second step of the fine level's recurrence: the post-smoothing kernel
  // of the V-cycle on its own, for measurement.
  void smooth_step_f32(const float* u, const float* b, float* y, int nrhs);
  // F, U: device, [nrhs][3][ldn] FP64; U holds x0 when use_x0.  Returns 0 on
  // convergence, 1 when maxiter was exhausted, 2 on a non-finite result.
  int solve(const double* F, double* U, int nrhs, double rtol, int maxiter, bool use_x0, int* iters_out);
  // x = M b; with dot_part the last fine-level smoothing kernel also writes the
  // per-block partials of <b, x> (apply grid, [nrhs][blocks]).
  // skip: the PCG scalar block whose done flags mark right-hand sides to leave
  // untouched (null outside a PCG inner loop: every rhs is smoothed)
  void vcycle(int l, const float* b, float* x, int nrhs, double* dot_part = nullptr, const double* skip = nullptr);
  // Tile shape of the FP32 apply: variant >= 0 sets it on every level,
  // variant < 0 times every shape on every multigrid level and keeps the
  // fastest per level (coarse levels are latency-bound and prefer short z
  // chunks).  us_out (may be null) receives the fine level's per-shape
  // microseconds.  Returns the fine level's shape.
  int autotune(int variant, float* us_out);
  // pin the FP32 apply tile shape of one multigrid level (what autotune picks
  // per level; captured V-cycles are dropped; the automatic sweep on the first
  // set_Ee is skipped from then on)
  void set_level_tile(int level, int variant);
  // The fine-level V-cycle z = M r on the PCG buffers, replayed from a hipGraph
  // captured once per rhs count (pointers and smoother coefficients are fixed;
  // Ee, dinv and the coarse inverse are read from fixed buffers).
  void vcycle_pcg(int nrhs);
  // Same on other buffers (x = M b), one graph per (nrhs, b, x).
  void vcycle_graph(const float* b, float* x, int nrhs);
  // Replay the V-cycle and the PCG iteration units from recorded hipGraphs.
  // Off by default: launching the same kernels one by one measured 1-2 %
  // faster at every grid size (the host stays ahead through the speculation,
  // and a graph launch left the device idle at each graph boundary).
  bool use_graphs = false;

  // Learned stage (a U-Net N on the node grid of multigrid level `level`,
  // z1 = gamma P N(R r) with the energy-optimal gamma = <r, z1> / <z1, K z1>):
  //   mode 0: off (plain multigrid PCG);
  //   mode 1: warm start: each inner solve starts from x = z1, r - K z1;
  //   mode 2: every iteration z = z1 + V(r - K z1) (two-stage, multiplicative),
  //           with flexible CG (beta = -<z, K p> / <p, K p>) because the stage
  //           is nonlinear in r.
  // gamma makes the network step a line search, so it never raises the energy
  // error; the FP64 true-residual refinement decides convergence as before.
  // e_ref scales the stiffness feature (cbrt(Ee / e_ref)) and the output (1 / e_ref).
  void set_learned(UNet* net, int level, double e_ref, int mode);
  int learned_mode() const { return net_ ? net_mode_ : 0; }
  int learned_level() const { return net_level_; }
  // z = P^l N(R^l r) on [nrhs][3][ldn] FP32 buffers.
  void learned_stage(const float* r, float* z, int nrhs);
  // Training pair for right-hand sides F and solutions U (FP64, [nrhs][3][ldn]):
  // in[b][V][8] the network input for r = F, tgt[b][V][3] the solution injected
  // onto the network level in output units.
  void learned_pair(const double* F, const double* U, int nrhs, int level, double e_ref, uint16_t* in, float* tgt);
  void invalidate_graphs() { drop_graphs(); }
  // Changes whenever recorded graphs of this context's kernels become stale
  // (tile shapes, smoother settings).
  uint64_t graph_generation() const { return graph_gen_; }
  // Levels l >= fuse_first_from form x1 = omega D^-1 b inside the first
  // Chebyshev kernel (one launch fewer per level and cycle)
  int fuse_first_from = 0;
  int tile() const { return lv_[0].tile; }

  const GridDims& grid() const { return lv_[0].g; }
  int levels() const { return (int)lv_.size(); }
  int64_t ldn() const { return lv_[0].g.ldn(); }
  const uint8_t* mask() const { return lv_[0].mask; }
  const double* ke() const { return ke_; }  // host, 576
  hipStream_t stream() const { return s_; }
  const double* Ee64() const { return Ee64_; }
  const KeCoef<double>& coef64() const { return k64_; }
  const Level& level(int l) const { return lv_[l]; }

  // Chebyshev smoother of degree nu_pre on [lmax / cheb_ratio, lmax] of D^-1 K.
  // lmax is the element bound lambda_max(ke) / ke_ii, an upper bound of
  // lambda_max(D^-1 K) for any stiffness scales (x^T K x = sum_e Ee x_e^T ke x_e
  // <= bound * sum_e Ee x_e^T diag(ke) x_e = bound * x^T D x), so the V-cycle
  // stays symmetric positive definite however heterogeneous Ee is.  Degree 3
  // on the fine level: 531 PCG iterations over the 100-iteration 192x96x96
  // schedule against 683 at degree 2 and 638 at degree 4, 5 % less time.
  int nu_pre = 3;
  // coarse Ee = power mean of the children, (mean E^q)^(1/q); takes effect at the next set_Ee
  double coarsen_power = 0.5;
  int nu_coarse = 3;  // degree on the coarse levels (l >= 1): cheaper per cell, fewer PCG iterations
  int nu_deep = 0;    // degree on levels l >= 2 when > 0 (there every dispatch is latency-bound)
  // When a solve needs more than deep_escalate_iters PCG iterations per right-hand
  // side (on average) while nu_deep is 0, the deep levels switch to degree
  // deep_escalate_degree for the rest of the context's life: a slowly converging
  // cycle (high stiffness contrast: keep-outs next to solid) gains from more
  // sweeps there, and they cost little.  0 disables it.
  int deep_escalate_iters = 20;
  int deep_escalate_degree = 13;  // large part: 185 / 161 / 150 / 136 / 137 ms per iteration at 3 / 5 / 7 / 13 / 15
  // applies the rule above after a solve; returns true when it switched
  bool maybe_escalate(const int* its, int nrhs);
  double cheb_ratio = 15.0;
  double cheb_lmax = 0.0;  // set from ke at construction
  float omega = 1.0f;      // default Jacobi weight in the apply arguments (the V-cycle sets its own)
  double inner_tol_floor = 1e-5;
  // the inner FP32 PCG stops at inner_tol_factor * rtol * |F| / |r| (relative to its start), so one
  // refinement pass usually meets rtol on the FP64 residual (0.85: 513 PCG
  // iterations on that schedule against 531 at 0.7)
  double inner_tol_factor = 0.85;
  int inner_max = 400;

 private:
  friend class BatchSolver;
  template <typename T>
  ApplyArgs<T> args(int l) const;
  void* alloc(size_t bytes);
  void alloc_solver(int nrhs_max);
  const OpEntry* optab(int l) const { return optab_.empty() ? nullptr : optab_[l]; }
  std::vector<FemContext*> members_;
  std::vector<OpEntry*> optab_;  // per level, one entry per member (device)
  int rpo_ = 1;

  hipStream_t s_;
  int nrhs_max_;
  std::vector<Level> lv_;
  std::vector<void*> owned_;
  double ke_[576];
  KeCoef<double> k64_{};
  double* Ee64_ = nullptr;
  double* ke_coarse_dev_ = nullptr;  // 576 doubles, coarsest h
  double* Ainv_ = nullptr;           // coarsest dense inverse
  double* Awork_ = nullptr;
  // PCG buffers
  float *px_ = nullptr, *pr_ = nullptr, *pz_ = nullptr, *pp_ = nullptr, *pq_ = nullptr;
  // The merged PCG iteration forms the direction inside the K p apply (apply
  // mode kApplyDir), which reads the old direction and writes the new one: the
  // two live in pp_ and pp2_, in turn by host slot (step with slot s writes
  // pdir(s)); a pass starts with its direction in pp_ and a step of slot 0.
  float* pp2_ = nullptr;
  float* pdir(int slot) const { return slot ? pp2_ : pp_; }
  double *x64_ = nullptr, *r64_ = nullptr;
  double *part_ = nullptr, *st_ = nullptr, *red_ = nullptr, *scale_ = nullptr;
  // pinned, 6 x 64 doubles: [0, 64) |r|^2 of the FP64 residual, [128, 192) |F|^2,
  // and the |r|^2 and done flag of a PCG step in one of two slots (rr_slot,
  // done_slot) that alternate between consecutive steps, so a step queued
  // before the host has read the previous step's values writes the other slot
  double* host_red_ = nullptr;
  double* host_red_dev_ = nullptr;  // its device alias: reductions are stored there directly
  static int rr_slot(int slot) { return slot ? 256 : 64; }
  static int done_slot(int slot) { return slot ? 320 : 192; }
  // pinned staging of the per-pass PCG set-up ([0, 64): scale per rhs, then the
  // 64 x kSt scalar blocks), read by one small kernel (upload) instead of two
  // copies from pageable memory; rewritten only after a host wait that follows
  // the upload of its previous contents
  double* host_up_ = nullptr;
  double* host_up_dev_ = nullptr;
  bool tuned_ = false;
  std::vector<int> tile_key() const;  // device and level dimensions
  void pick_tiles();                  // cached shapes for this hierarchy, else the timing sweep
  bool coarse_stale_ = true, coarse_valid_ = false;
  hipStream_t cap_s_ = nullptr;             // private stream used only to record graphs
  hipEvent_t ev_ = nullptr, ev2_ = nullptr;  // mark the residual-norm copies of PCG steps
  std::vector<hipGraphExec_t> vgraph_;      // per rhs count
  std::map<std::tuple<int, const void*, void*>, hipGraphExec_t> vgraph_bx_;
  // the two halves of a plain PCG iteration, per rhs count: step (K p with
  // <p, K p>, alpha, x and r update with |r|^2 to the host) and direction
  // (V-cycle with <r, z>, beta), and both as one unit (direction of step k,
  // then step k + 1, whose apply forms p = z + beta p): one launch per
  // iteration once the loop runs.  With graphs on, the graphs holding a step
  // are recorded per host slot: [2 * nrhs + slot]
  std::vector<hipGraphExec_t> sgraph_, dgraph_, cgraph_;
  void pcg_step(int nrhs, int slot);
  void pcg_direction(int nrhs);
  void pcg_direction_step(int nrhs, int slot);
  // fused: p = z + beta p formed by the apply (after issue_direction), else p
  // is already in pdir(slot)
  void issue_step(int nrhs, int slot, bool fused);
  void issue_direction(int nrhs);  // V-cycle, <r, z>, beta (the direction itself is the next step's)
  // records what issue() launches on s_ (redirected to the capture stream)
  template <typename F>
  hipGraphExec_t capture(F&& issue);
  void drop_graphs();
  uint64_t graph_gen_ = 0;
  // pz_ = M pr_; returns true when part_ already holds the apply-grid partials of <pr_, pz_>
  bool precondition(int nrhs);
  void scaled_network_step(const float* r, float* z1, float* t, int nrhs);
  void learned_warm_start(int nrhs);
  UNet* net_ = nullptr;
  int net_level_ = 1, net_mode_ = 0;
  double net_eref_ = 1.0;
  uint16_t* net_in_ = nullptr;
  float* net_out_ = nullptr;
  double* net_nrm_ = nullptr;
  float *pz1_ = nullptr, *pt_ = nullptr;
};

}  // namespace anneal
