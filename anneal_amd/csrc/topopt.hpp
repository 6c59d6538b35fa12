// Density-side kernels of the SIMP loop: hat filter, Heaviside projection,
// sensitivities and the OC bisection, all FP64 on the full voxel grid.
#pragma once
#include <vector>

#include "stencil.hpp"

namespace anneal {

// Element flags
constexpr uint8_t kActive = 1, kFrozen = 2;

// OC bisection state on the device (FP64).  kOcLam is the multiplier the next
// trial-volume kernel evaluates; (lo, hi) is the reference's bisection state;
// [A, B] is a certified bracket of the root of vol(lam) = vt (FA, FB: vol - vt
// there, TA, TB: their logs); (T1, F1), (T2, F2) are log(lam) and vol - vt of the last two
// evaluations; Slope is d(vol)/d(log lam) carried over from the previous
// bisection (0: none), SlopeNext the one this bisection hands on; Grow is the
// log step taken from a one-sided first evaluation when no slope is known.
constexpr int kOcLo = 0, kOcHi = 1, kOcLam = 2, kOcLamLast = 3, kOcDone = 4, kOcSteps = 5, kOcVt = 6, kOcVol = 7,
              kOcA = 8, kOcFA = 9, kOcB = 10, kOcFB = 11, kOcHasA = 12, kOcHasB = 13, kOcSide = 14, kOcEvals = 15,
              kOcGrow = 16, kOcT1 = 17, kOcF1 = 18, kOcT2 = 19, kOcF2 = 20, kOcSlope = 21, kOcSlopeNext = 22,
              kOcTA = 23, kOcTB = 24;  // log(A), log(B)
constexpr int kOcLen = 26;

struct FilterArgs;

class FilterContext {
 public:
  // Linear hat kernel max(0, r - |d|) on [-R, R]^3, R = max(ceil(r) - 1, 1)
  // (anneal/optimize.py:24-34), normalised by Hs = conv(active) with 0 -> 1.
  FilterContext(int nx, int ny, int nz, double r_vox, const uint8_t* flags_host, hipStream_t stream);
  ~FilterContext();
  FilterContext(const FilterContext&) = delete;
  FilterContext& operator=(const FilterContext&) = delete;

  void forward(const double* x, double* y);   // optimize.py:41-43
  void backward(const double* d, double* y);  // optimize.py:45-47
  // backward(d1, y1) and backward(d2, y2) in one pass over the grid; with
  // stats_dev, also stats_dev[3] = design count, [4] = sum of -y1, [5] = sum of
  // y2 over design voxels (active, not frozen) and [0..2] = 0, the layout of
  // design_stats.
  void backward_pair(const double* d1, double* y1, const double* d2, double* y2, double* stats_dev);
  // xPhys, dxp = project(forward(x)), frozen -> 1 (optimize.py:81-93).
  // Also returns sum(phi * xPhys) over active voxels in *vol_dev (device).
  void physical(const double* x, double beta, const double* phi, double* xphys, double* dxp, double* vol_dev);
  // The geometric bisection of optimize.py:144-161 on the device, replayed
  // exactly (same trial multipliers, decisions and stop) from fewer volume
  // evaluations: vol(lam) is non-increasing, so a certified bracket [A, B] of
  // its root decides every trial outside it.  The multipliers evaluated follow
  // a safeguarded secant in log(lam): the first at hint > 0 (a guess of the
  // root, e.g. the previous iteration's), the second a Newton step with the
  // slope the previous bisection of this context measured (or a log step of
  // width, <= 0: 0.005, toward the root), then secant steps through the last
  // two evaluations kept inside [A, B]; an estimate within 1e-12 of the last
  // evaluation is stepped 1e-12 past it, so the bracket closes on the root.
  // xnew receives the last trial design; returns the number of trials (the
  // reference's count).
  int oc_bisect(const double* x, const double* dcx, const double* dvx, const double* phi, double beta, double move,
                double lo, double hi, double vtarget, double* xnew, double hint = 0.0, double width = 0.0,
                int* evals = nullptr, double* lam_last = nullptr);
  // oc_bisect in three calls, so that one host thread can keep several
  // contexts' bisections in flight: oc_begin queues the first evaluations and
  // the state read-back; oc_poll waits for that read-back only, returns true
  // once the bisection has stopped and otherwise queues the next evaluations;
  // oc_end queues the final design and returns the trial count.  oc_bisect is
  // these calls in sequence (same launches).
  void oc_begin(const double* x, const double* dcx, const double* dvx, const double* phi, double beta, double move,
                double lo, double hi, double vtarget, double* xnew, double hint = 0.0, double width = 0.0);
  // oc_begin with the bracket taken on the device from design statistics
  // (design_stats: stats[3] design count, stats[4] sum -dcx, stats[5] sum dvx),
  // as optimize.py:146-147 forms it, and the first evaluation at ratio times
  // its scale (ratio <= 0: none): no host wait for the statistics.
  void oc_begin_stats(const double* x, const double* dcx, const double* dvx, const double* phi, double beta,
                      double move, const double* stats, double ratio, double vtarget, double* xnew,
                      double width = 0.0);
  bool oc_poll();
  int oc_end(int* evals = nullptr, double* lam_last = nullptr);

  const GridDims& grid() const { return g_; }
  const uint8_t* flags() const { return flags_; }
  const double* Hs() const { return Hs_; }
  int R() const { return R_; }
  hipStream_t stream() const { return s_; }

 private:
  GridDims g_;
  int R_;
  std::vector<double> w_;  // hat weight per squared voxel distance 0 .. 3R^2
  hipStream_t s_;
  uint8_t* flags_ = nullptr;
  double *Hs_ = nullptr, *invHs_ = nullptr, *r0_ = nullptr, *part_ = nullptr, *oc_ = nullptr, *host_ = nullptr;
  double* host_dev_ = nullptr;  // device alias of host_ (pinned): state read-back and upload without copies
  double* wt_ = nullptr;        // w_ in device memory
  double* stg_ = nullptr;       // staged values of the gather path (radii above the ring kernels')
  int nblk_ = 0;
  // the bisection in flight (oc_begin .. oc_end)
  struct OcRun {
    const double *x, *phi;
    double* xnew;
    const double *dcx, *dvx;
    double beta, move;
    int launched;
    bool active;
    double* state;  // the state after the last batch (one of the two buffers in oc_)
  } run_{};
  double oc_slope_ = 0.0;  // d(vol)/d(log lam) measured by the last bisection (0: none yet)
  hipEvent_t oc_ready_ = nullptr;  // recorded after each state read-back
  void oc_start(const double* x, const double* dcx, const double* dvx, const double* phi, double beta, double move,
                double lo, double hi, double vtarget, double* xnew, double hint, double width, const double* stats,
                double ratio);
  void oc_launch(int count);
  FilterArgs args() const;  // launch arguments common to every mode
};

// ce_r(e) = u_e^T ke u_e for E = 1 (anneal/fem.py:239-242), out [nrhs][nelem].
void element_energy(const double* U, int nrhs, int64_t ldn, const GridDims& g, const KeCoef<double>& k,
                    const uint8_t* flags, double* out, hipStream_t s);
// Strain-displacement matrix at the element centre, elasticity matrix and the
// corner order of the 24 element dofs (anneal/fem.py:20-21, 36-84).
struct StressMats {
  double B0[6][24];
  double D[6][6];
  int local[8][3];
};
// sig[r][e][6] = E_e D B0 u_e (Voigt xx yy zz xy yz zx), vm[r][e] its von Mises
// stress (fem.py:244-263); E per element (full grid) or the scalar Es when E is
// null; inactive voxels get zeros.  vm may be null.
void element_stress(const double* U, int nrhs, int64_t ldn, const GridDims& g, const StressMats& m,
                    const uint8_t* flags, const double* E, double Es, double* sig, double* vm, hipStream_t s);
// dc*dxp and phi*dxp (anneal/optimize.py:118, 125-142):
//   dce_r = -p x^(p-1) (1-emin) E0 phi ce_r; dc = min(sum_r coef_r dce_r, 0)
void sensitivity(const double* U, int nrhs, int64_t ldn, const GridDims& g, const KeCoef<double>& k,
                 const uint8_t* flags, const double* xphys, const double* dxp, const double* phi, const double* coef,
                 double p, double emin, double E0, double* dcdx, double* dvdx, hipStream_t s);
// Ee = E0 phi (emin + x^p (1 - emin)) on active voxels, 0 elsewhere (optimize.py:118).
void stiffness_law(const double* xphys, const double* phi, const uint8_t* flags, double E0, double emin, double p,
                   int64_t n, double* Ee, hipStream_t s);
// Per-iteration statistics over design voxels (active, not frozen):
//   out[0] = sum phi*xphys (active), out[1] = #grey (0.1 < xphys < 0.9), out[2] = max |xnew - xold|,
//   out[3] = #design, out[4] = sum(-dcx) over design, out[5] = sum(dvx) over design
void design_stats(const double* xold, const double* xnew, const double* xphys, const double* phi, const double* dcx,
                  const double* dvx, const uint8_t* flags, int64_t n, double* out_dev, double* part, hipStream_t s);

}  // namespace anneal
