// Mixed-precision PCG over several independent FEM contexts together (grids of
// different sizes, each with its own block of right-hand sides).
//
// One inner PCG iteration of a problem is launched on the problem's own
// stream (or replayed from a hipGraph recorded once per rhs count and host
// slot, when the context's graph option is on), forked from the launch
// stream once per pass and joined back at its end, so independent
// problems run concurrently and finished ones cost nothing.  Each problem's
// next iteration is queued before the host waits for its current one, so the
// device does not idle while the host reads the norms.  A right-hand side
// whose step meets its inner tolerance is marked done on the device: the
// kernels queued after it skip its workgroups, so its x, r and p stay
// untouched and it costs no bandwidth while the other right-hand sides keep
// iterating.  The FP64 true-residual refinement and its decisions are the
// same as in FemContext::solve.
#include "batch.hpp"

#include <algorithm>
#include <cmath>

namespace anneal {

BatchSolver::BatchSolver(const std::vector<FemContext*>& fems) : f_(fems) {
  ANNEAL_CHECK(!f_.empty() && f_.size() <= 64, "a batch holds 1..64 FEM contexts");
  s_ = f_[0]->s_;
  for (FemContext* f : f_) {
    ANNEAL_CHECK(f != nullptr, "null FEM context in the batch");
    ANNEAL_CHECK(f->s_ == s_, "the contexts of a batch must share one stream");
  }
  br_.resize(f_.size());
  graph_.assign(3 * f_.size(), nullptr);
  graph_nrhs_.assign(f_.size(), 0);
  graph_gen_.assign(f_.size(), 0);
  join_.resize(f_.size());
  ev_.resize(2 * f_.size());
  for (size_t i = 0; i < f_.size(); ++i) {
    ANNEAL_HIP(hipStreamCreateWithFlags(&br_[i], hipStreamNonBlocking));
    ANNEAL_HIP(hipEventCreateWithFlags(&join_[i], hipEventDisableTiming));
    for (int k = 0; k < 2; ++k) ANNEAL_HIP(hipEventCreateWithFlags(&ev_[2 * i + k], hipEventDisableTiming));
  }
  ANNEAL_HIP(hipEventCreateWithFlags(&fork_, hipEventDisableTiming));
}

BatchSolver::~BatchSolver() {
  (void)hipStreamSynchronize(s_);
  for (auto s : br_) (void)hipStreamSynchronize(s);
  drop_graphs();
  for (auto s : br_) (void)hipStreamDestroy(s);
  for (auto e : join_) (void)hipEventDestroy(e);
  for (auto e : ev_)
    if (e) (void)hipEventDestroy(e);
  if (fork_) (void)hipEventDestroy(fork_);
}

void BatchSolver::drop_graphs() {
  for (auto& g : graph_)
    if (g) {
      (void)hipGraphExecDestroy(g);
      g = nullptr;
    }
}

// Unit k of a problem's inner PCG, issued on its current stream: unit 0 is
// the pass's first step (the set-up formed its direction); a later unit is the
// direction of step k - 1 (V-cycle with <r, z>, beta) followed by step k, whose
// apply forms p = z + beta p (FemContext::solve's merged iteration).  Step k
// stores its norms in host slot k % 2 and its direction in pdir(k % 2).
void BatchSolver::issue_unit(FemContext& f, int nrhs, int k) {
  const GridDims& g0 = f.lv_[0].g;
  const int64_t n = g0.nnode(), ld = g0.ldn();
  const int nblk_apply = apply_blocks(g0.nx, g0.ny, g0.nz, f.lv_[0].tile);
  const int slot = k & 1;
  if (k > 0) {
    // the cycle's last fine smoothing kernel forms the <r, z> partials (a
    // one-level hierarchy is a dense solve only: separate dot)
    const bool fused = f.lv_.size() > 1;
    f.vcycle(0, f.pr_, f.pz_, nrhs, fused ? f.part_ : nullptr, f.st_);
    if (!fused) dot_partial(f.pr_, f.pz_, n, ld, nrhs, f.part_, f.s_, f.st_);
    finalize_st(f.part_, fused ? nblk_apply : kDotBlocks, nrhs, f.st_, kRZnew, 2, f.s_);
  }
  ApplyArgs<float> af = f.args<float>(0);
  af.y = f.pq_;
  af.dot_part = f.part_;
  af.skip = f.st_;
  if (k > 0) {
    af.u = f.pdir(slot ^ 1);
    af.b = f.pz_;
    af.pout = f.pdir(slot);
    launch_apply(af, kApplyDir, nrhs, f.s_);
  } else {
    af.u = f.pdir(0);
    launch_apply(af, kApply, nrhs, f.s_);
  }
  finalize_st(f.part_, nblk_apply, nrhs, f.st_, kPQ, 1, f.s_);
  // |r|^2 to the host; a rhs that meets its inner tolerance is marked done on the device
  pcg_update_converge(f.px_, f.pr_, f.pdir(slot), f.pq_, f.st_, n, ld, nrhs, f.part_, f.s_,
                      f.host_red_dev_ + FemContext::rr_slot(slot), f.host_red_dev_ + FemContext::done_slot(slot));
}

// The recorded variants of a problem's units: 0 the first step, 1 + s a later
// unit whose step has host slot s.
static int unit_variant(int k) { return k == 0 ? 0 : 1 + (k & 1); }

void BatchSolver::record(int p, int nrhs) {
  FemContext& f = *f_[p];
  if (graph_[3 * p] && graph_nrhs_[p] == nrhs && graph_gen_[p] == f.graph_generation()) return;
  ANNEAL_HIP(hipStreamSynchronize(br_[p]));
  for (int k : {0, 1, 2}) {  // one unit of each variant
    hipGraphExec_t& ge = graph_[3 * p + unit_variant(k)];
    if (ge) {
      (void)hipGraphExecDestroy(ge);
      ge = nullptr;
    }
    hipStream_t saved = f.s_;
    hipGraph_t g = nullptr;
    try {
      ANNEAL_HIP(hipStreamBeginCapture(br_[p], hipStreamCaptureModeThreadLocal));
      f.s_ = br_[p];
      issue_unit(f, nrhs, k);
      f.s_ = saved;
      ANNEAL_HIP(hipStreamEndCapture(br_[p], &g));
    } catch (...) {
      f.s_ = saved;
      hipGraph_t dummy = nullptr;
      (void)hipStreamEndCapture(br_[p], &dummy);
      if (dummy) (void)hipGraphDestroy(dummy);
      throw;
    }
    const hipError_t e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    ANNEAL_HIP(e);
  }
  graph_nrhs_[p] = nrhs;
  graph_gen_[p] = f.graph_generation();
}

void BatchSolver::solve(const std::vector<const double*>& F, const std::vector<double*>& U,
                        const std::vector<int>& nrhs, double rtol, int maxiter, bool use_x0, int* status,
                        int* iters /* [nprob][64] */) {
  const int np = (int)f_.size();
  ANNEAL_CHECK((int)F.size() == np && (int)U.size() == np && (int)nrhs.size() == np, "one F, U, nrhs per problem");
  struct St {
    std::vector<double> fnorm, rnorm, tol2, sc, hst;
    std::vector<int> its;
    std::vector<char> done, idone;
    int status = 1, total = 0;
    bool finished = false, bad = false, inner = false;
  };
  std::vector<St> S(np);
  for (int p = 0; p < np; ++p) {
    FemContext& f = *f_[p];
    const int nr = nrhs[p];
    ANNEAL_CHECK(nr >= 1 && nr <= f.nrhs_max_ && nr <= 64, "nrhs out of range for a batch member");
    f.set_rhs_count(nr);
    ANNEAL_CHECK(f.learned_mode() == 0, "batched solves use the plain multigrid preconditioner");
    if (!f.coarse_valid_ || (f.coarse_stale_ && f.rebuild_each_solve)) f.rebuild_coarse();
    const GridDims& g0 = f.lv_[0].g;
    if (!use_x0) ANNEAL_HIP(hipMemsetAsync(U[p], 0, (size_t)nr * 3 * g0.ldn() * sizeof(double), s_));
    dot(F[p], F[p], g0.nnode(), g0.ldn(), nr, f.part_, f.red_, s_, f.host_red_dev_);
    St& st = S[p];
    st.fnorm.assign(nr, 0.0);
    st.rnorm.assign(nr, 0.0);
    st.tol2.assign(nr, 0.0);
    st.sc.assign(nr, 0.0);
    st.hst.assign((size_t)nr * kSt, 0.0);
    st.its.assign(nr, 0);
    st.done.assign(nr, 0);
    st.idone.assign(nr, 0);
  }
  ANNEAL_HIP(hipStreamSynchronize(s_));
  for (int p = 0; p < np; ++p)
    for (int r = 0; r < nrhs[p]; ++r) S[p].fnorm[r] = std::sqrt(f_[p]->host_red_[r]);

  for (int outer = 0; outer < 50; ++outer) {
    // FP64 true residuals of every unfinished problem, one wait
    for (int p = 0; p < np; ++p) {
      if (S[p].finished) continue;
      FemContext& f = *f_[p];
      const GridDims& g0 = f.lv_[0].g;
      ApplyArgs<double> a = f.args<double>(0);
      a.u = U[p];
      a.b = F[p];
      a.y = f.r64_;
      a.dot_part = f.part_;  // |r|^2 partials from the residual kernel itself
      launch_apply(a, kResid, nrhs[p], s_);
      finalize(f.part_, apply_blocks(g0.nx, g0.ny, g0.nz, kApplyVariant64), nrhs[p], f.red_, s_, f.host_red_dev_);
    }
    ANNEAL_HIP(hipStreamSynchronize(s_));
    std::vector<int> active;
    for (int p = 0; p < np; ++p) {
      St& st = S[p];
      st.inner = false;
      if (st.finished) continue;
      FemContext& f = *f_[p];
      const int nr = nrhs[p];
      bool all = true;
      for (int r = 0; r < nr; ++r) {
        st.rnorm[r] = std::sqrt(f.host_red_[r]);
        if (!std::isfinite(st.rnorm[r])) st.bad = true;
        st.done[r] = (st.rnorm[r] <= rtol * st.fnorm[r]) ? 1 : 0;
        all = all && st.done[r];
      }
      if (st.bad) {
        st.status = 2;
        st.finished = true;
        continue;
      }
      if (all) {
        st.status = 0;
        st.finished = true;
        continue;
      }
      if (st.total >= maxiter) {
        st.status = 1;
        st.finished = true;
        continue;
      }
      // inner FP32 PCG set-up (as in FemContext::solve)
      const GridDims& g0 = f.lv_[0].g;
      const int64_t n = g0.nnode(), ld = g0.ldn();
      for (int r = 0; r < nr; ++r) {
        st.sc[r] = st.done[r] ? 0.0 : 1.0 / st.rnorm[r];
        const double want = st.done[r] ? 1.0 : f.inner_tol_factor * rtol * st.fnorm[r] / st.rnorm[r];
        const double t = std::max(want, f.inner_tol_floor);
        st.tol2[r] = t * t;
        for (int k = 0; k < kSt; ++k) st.hst[r * kSt + k] = 0.0;
        st.hst[r * kSt + kDone] = st.done[r] ? 1.0 : 0.0;
        st.hst[r * kSt + kTol2] = st.tol2[r];
        st.idone[r] = st.done[r];
      }
      std::copy(st.sc.begin(), st.sc.begin() + nr, f.host_up_);
      std::copy(st.hst.begin(), st.hst.end(), f.host_up_ + 64);
      upload(f.host_up_dev_, f.scale_, nr, f.host_up_dev_ + 64, f.st_, (int)st.hst.size(), s_);
      to_f32_scaled(f.r64_, f.pr_, f.scale_, n, ld, nr, s_);
      ANNEAL_HIP(hipMemsetAsync(f.px_, 0, (size_t)nr * 3 * ld * sizeof(float), s_));
      f.vcycle_pcg(nr);  // skips the rhs that already meet the tolerance
      if (f.lv_.size() > 1) {
        finalize_st(f.part_, apply_blocks(g0.nx, g0.ny, g0.nz, f.lv_[0].tile), nr, f.st_, kRZ, 0, s_);
      } else {
        dot_partial(f.pr_, f.pz_, n, ld, nr, f.part_, s_, f.st_);
        finalize_st(f.part_, kDotBlocks, nr, f.st_, kRZ, 0, s_);
      }
      pcg_dir(f.pp_, f.pz_, f.st_, n, ld, nr, 1, s_);
      st.inner = true;
      active.push_back(p);
    }
    if (active.empty()) break;
    for (int p : active)
      if (f_[p]->use_graphs) record(p, nrhs[p]);
    // Every problem iterates on its own stream from the set-up above (one fork
    // per pass), and its iteration k + 1 is launched before the host waits for
    // its iteration k, so the device does not idle on the host's round trip.
    // Iteration k stores its norms in host slot k % 2.  The step that meets a
    // rhs's tolerance marks it done on the device, so the kernels of an
    // iteration queued after it skip that rhs; an iteration queued for a
    // problem whose every rhs is already done costs its launches only.  No
    // iteration is launched past inner_max or maxiter.
    ANNEAL_HIP(hipEventRecord(fork_, s_));
    for (int p : active) ANNEAL_HIP(hipStreamWaitEvent(br_[p], fork_, 0));
    std::vector<int> launched(np, 0), counted(np, 0);  // iterations of this pass, per problem
    auto launch_next = [&](int p) {
      St& st = S[p];
      const int k = launched[p];
      if (!st.inner || k >= f_[p]->inner_max || st.total + (k - counted[p]) >= maxiter) return;
      FemContext& f = *f_[p];
      if (f.use_graphs) {
        ANNEAL_HIP(hipGraphLaunch(graph_[3 * p + unit_variant(k)], br_[p]));
      } else {  // the same kernels launched one by one on the problem's stream
        hipStream_t saved = f.s_;
        f.s_ = br_[p];
        try {
          issue_unit(f, nrhs[p], k);
        } catch (...) {
          f.s_ = saved;
          throw;
        }
        f.s_ = saved;
      }
      ANNEAL_HIP(hipEventRecord(ev_[2 * p + (k & 1)], br_[p]));
      ++launched[p];
    };
    for (int p : active) launch_next(p);
    for (int k = 0;; ++k) {
      for (int p : active) launch_next(p);  // iteration k + 1, queued ahead
      bool any = false;
      for (int p : active) {
        St& st = S[p];
        if (!st.inner || launched[p] <= k) continue;  // no iteration k of this pass
        any = true;
        FemContext& f = *f_[p];
        ANNEAL_HIP(hipEventSynchronize(ev_[2 * p + (k & 1)]));
        ++counted[p];
        ++st.total;
        bool iall = true;
        for (int r = 0; r < nrhs[p]; ++r) {
          if (st.idone[r]) continue;
          ++st.its[r];
          const double rr = f.host_red_[FemContext::rr_slot(k & 1) + r];
          if (!std::isfinite(rr)) st.bad = true;
          if (f.host_red_[FemContext::done_slot(k & 1) + r] != 0.0) st.idone[r] = 1;  // |r|^2 <= tol2
          iall = iall && st.idone[r];
        }
        // this problem's inner solve ends: no further iteration is launched
        if (st.bad || iall || st.total >= maxiter || k + 1 >= f.inner_max) st.inner = false;
      }
      if (!any) break;
    }
    for (int p : active) {  // the launch stream waits for every iteration queued, counted or not
      ANNEAL_HIP(hipEventRecord(join_[p], br_[p]));
      ANNEAL_HIP(hipStreamWaitEvent(s_, join_[p], 0));
    }
    for (int p : active) {
      St& st = S[p];
      FemContext& f = *f_[p];
      if (st.bad) {
        st.status = 2;
        st.finished = true;
        continue;
      }
      const GridDims& g0 = f.lv_[0].g;
      for (int r = 0; r < nrhs[p]; ++r) f.host_up_[r] = st.done[r] ? 0.0 : st.rnorm[r];
      upload(f.host_up_dev_, f.scale_, nrhs[p], nullptr, nullptr, 0, s_);
      axpy_f32_to_f64(f.px_, U[p], f.scale_, g0.nnode(), g0.ldn(), nrhs[p], s_);
    }
  }
  ANNEAL_HIP(hipStreamSynchronize(s_));
  for (int p = 0; p < np; ++p) {
    status[p] = S[p].finished ? S[p].status : 1;
    for (int r = 0; r < 64; ++r) iters[p * 64 + r] = r < nrhs[p] ? S[p].its[r] : 0;
    f_[p]->maybe_escalate(S[p].its.data(), nrhs[p]);  // a switch bumps the graph generation: re-recorded
  }
}

}  // namespace anneal
