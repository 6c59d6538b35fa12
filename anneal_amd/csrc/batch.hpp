// Several independent FEM problems (any grid sizes) solved together.
#pragma once
#include <vector>

#include "fem_ctx.hpp"

namespace anneal {

class BatchSolver {
 public:
  // The contexts must share one stream; they stay owned by the caller.
  explicit BatchSolver(const std::vector<FemContext*>& fems);
  ~BatchSolver();
  BatchSolver(const BatchSolver&) = delete;
  BatchSolver& operator=(const BatchSolver&) = delete;

  // Per problem p: F[p], U[p] FP64 [nrhs[p]][3][ldn_p] (U holds x0 when
  // use_x0).  status[p]: 0 converged, 1 maxiter exhausted, 2 non-finite;
  // iters[p * 64 + r]: PCG iterations of rhs r.
  void solve(const std::vector<const double*>& F, const std::vector<double*>& U, const std::vector<int>& nrhs,
             double rtol, int maxiter, bool use_x0, int* status, int* iters);
  int size() const { return (int)f_.size(); }

 private:
  // unit k of a problem's inner PCG (direction of step k - 1, then step k)
  void issue_unit(FemContext& f, int nrhs, int k);
  void record(int p, int nrhs);  // problem p's unit graphs: the first step, and a later unit per host slot
  void drop_graphs();

  std::vector<FemContext*> f_;
  hipStream_t s_ = nullptr;
  std::vector<hipStream_t> br_;  // one stream per problem
  std::vector<hipEvent_t> join_;
  std::vector<hipEvent_t> ev_;  // [2 * p + slot]: end of problem p's last iteration in that slot
  hipEvent_t fork_ = nullptr;
  std::vector<hipGraphExec_t> graph_;  // [3 * p + variant]
  std::vector<int> graph_nrhs_;
  std::vector<uint64_t> graph_gen_;
};

}  // namespace anneal
