// Density filter, projection, sensitivities and OC bisection (FP64, gfx950).
//
// The filter is a direct stencil, marched along z (see k_filter): each input
// layer is staged in LDS once and scattered into register accumulators of the
// output layers it reaches; the staged value is computed on the fly (masked
// design, scaled sensitivity or the OC trial design).  The hat kernel
// max(0, r-|d|) is not separable (its unfoldings have rank 2-3), so the direct
// form is the exact one.
#include "topopt.hpp"

#include <algorithm>
#include <cmath>
#include <type_traits>
#include <vector>

namespace anneal {

// kBwd2: two backward passes in one launch (in -> out, in2 -> out2) with the
// design statistics of the pair (count, sum of -out, sum of out2) as partials.
enum FilterMode : int { kHs = 0, kFwd = 1, kBwd = 2, kPhys = 3, kOc = 4, kBwd2 = 5 };

// Radii R <= kRingMaxR run the z-marching ring-sum kernel (k_filter<R, MODE>);
// larger ones stage the filtered quantity in a scratch grid and gather each
// output's (2R+1)^3 cube from it (k_filter_stage / k_filter_gather).
constexpr int kRingMaxR = 5;
constexpr int kWc = 3 * kRingMaxR * kRingMaxR + 1;  // squared distances 0 .. 3R^2

struct FilterArgs {
  const double* in;
  const double* in2;  // kBwd2: the second input
  const double* dcx;
  const double* dvx;
  const uint8_t* flags;
  const double* invHs;
  const double* phi;
  double* out;
  double* out2;
  double* part;
  const double* oc;       // kOc: bisection state before this launch's step
  double* oc_out;         // kOc: state after it (written by workgroup 0)
  const double* part_in;  // kOc with prev: volume partials of the previous evaluation, not yet in oc
  int prev;
  double wc[kWc];    // hat weight max(0, r - sqrt(d2)) for squared distance d2 = 0..3R^2 (R <= kRingMaxR)
  const double* wt;  // the same weights in device memory (any R; the gather kernel reads them)
  double* stg;       // staged values of the gather path (nelem doubles)
  int R;
  double beta, move;
  double pt0, pinv;  // projection constants tanh(beta eta) and 1 / (tanh(beta eta) + tanh(beta (1 - eta))) (set_beta)
  int nx, ny, nz;
};

// The projection's two element-independent tanh values are formed once on the
// host, in the order optimize.py:85-87 uses.
static void set_beta(FilterArgs& a, double beta) {
  const double eta = 0.5;
  a.beta = beta;
  a.pt0 = beta > 0.0 ? std::tanh(beta * eta) : 0.0;
  a.pinv = beta > 0.0 ? 1.0 / (a.pt0 + std::tanh(beta * (1.0 - eta))) : 1.0;
}

constexpr int FTX = 32, FTY = 8, FZC = 8;

// tanh(y) = sign(y) (1 - e) / (1 + e) with e = exp(-2|y|), written with
// m = expm1(-2|y|) in (-1, 0] as -m / (2 + m) so that small |y| keeps its
// relative accuracy: within 3 ulp of the C library's tanh over |y| <= 40 and
// at 1e-3 scale, exact 0 and +-1 at the ends; one expm1 and one division
// instead of the library tanh's longer evaluation (this is the per-element
// cost of every projected filter launch).
__device__ __forceinline__ double tanh_expm1(double y) {
  const double m = expm1(-2.0 * fabs(y));
  return copysign(-m / (2.0 + m), y);
}

// Heaviside projection of optimize.py:81-87 (eta = 1/2) and its derivative;
// pt0 = tanh(beta eta) and pinv = 1 / (pt0 + tanh(beta (1 - eta))) come from
// set_beta (a multiply by pinv instead of the division: within 1 ulp more).
__device__ __forceinline__ void project(double xt, const FilterArgs& a, double& xp, double& dxp) {
  if (a.beta <= 0.0) {
    xp = xt;
    dxp = 1.0;
    return;
  }
  const double th = tanh_expm1(a.beta * (xt - 0.5));
  xp = (a.pt0 + th) * a.pinv;
  dxp = a.beta * (1.0 - th * th) * a.pinv;
}

// z-marching filter.  A workgroup owns an FTX x FTY tile of element columns
// and a chunk of FZC output layers.  It walks the input layers Z0-R ..
// Z0+FZC-1+R once: each layer (tile plus an R-wide halo in x and y) is
// written to one of two LDS buffers, and every thread sums the 25 (or 9)
// values of its (dy, dx) window into rings of equal dy^2 + dx^2.  The hat
// weight depends only on dz^2 + dy^2 + dx^2, so the layer's contribution to
// the output layer |dz| away is sum_m w(dz^2 + m) * ring_m: 6 rings and 3
// distinct |dz| for R = 2 (18 FMAs per input layer instead of 93).  A window
// of 2R+1 register accumulators follows the march; output layer j is complete
// after input layer j + 2R and is finished there (normalisation, projection,
// volume).  Input layers are fetched two steps ahead in registers, the
// epilogue operands of a layer one step ahead.  The summation order of every
// output is fixed (input layers in +z; rings in order; taps in dy, dx order).

// Sum the volume partials in a fixed order, record the evaluation in the
// bracket, replay every bisection trial it decides, and choose the next
// multiplier to evaluate.
constexpr double kOcTau = 1e-12;       // log-step past the root estimate that closes the bracket
// Evaluations queued by oc_begin and by each oc_poll that finds the bisection
// still running (launches after the stop only carry the state forward, ~4 us
// each): the replay needs 5-6 evaluations on most SIMP iterations.
constexpr int kOcFirstBatch = 7, kOcNextBatch = 4;
constexpr double kOcGrowStep = 0.005;  // default log step from a one-sided first evaluation

__device__ void oc_record(double* oc, double lam, double t, double f) {
  if (f > 0.0) {
    if (oc[kOcHasA] == 0.0 || lam > oc[kOcA]) {
      oc[kOcA] = lam;
      oc[kOcTA] = t;
      oc[kOcFA] = f;
      oc[kOcHasA] = 1.0;
      if (oc[kOcSide] > 0.0 && oc[kOcHasB] != 0.0) oc[kOcFB] *= 0.5;  // Illinois
    }
    oc[kOcSide] = 1.0;
  } else {
    if (oc[kOcHasB] == 0.0 || lam < oc[kOcB]) {
      oc[kOcB] = lam;
      oc[kOcTB] = t;
      oc[kOcFB] = f;
      oc[kOcHasB] = 1.0;
      if (oc[kOcSide] < 0.0 && oc[kOcHasA] != 0.0) oc[kOcFA] *= 0.5;
    }
    oc[kOcSide] = -1.0;
  }
}

// Fixed-order sum of nblk volume partials over a 256-thread workgroup (thread
// 0 holds the result): every caller gets the same bits from the same partials.
__device__ double oc_volume(const double* __restrict__ part, int nblk, double* sh) {
  // partials threadIdx.x + 256 i in increasing i; four loads in flight at a time
  // (clamped addresses, the padding added as 0)
  double s = 0.0;
  for (int i0 = threadIdx.x; i0 < nblk; i0 += 4 * 256) {
    double v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) v[u] = part[min(i0 + u * 256, nblk - 1)];
#pragma unroll
    for (int u = 0; u < 4; ++u) s += i0 + u * 256 < nblk ? v[u] : 0.0;
  }
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  return sh[0] + sh[1] + sh[2] + sh[3];
}

// The multiplier to evaluate next when the reference's trial lam lies inside
// the bracket (or on the unknown side).  Only which multipliers are evaluated
// depends on this choice, never a decision or the result.  One thread of
// every workgroup runs it before the trial starts, so it takes no logarithm
// (those of the evaluated multipliers are kept in the state) and one exp.
__device__ double oc_choose(const double* oc, double lo, double hi, double lam) {
  const bool hasA = oc[kOcHasA] != 0.0, hasB = oc[kOcHasB] != 0.0;
  const double n = oc[kOcEvals];
  const double t1 = oc[kOcT1], f1 = oc[kOcF1], t2 = oc[kOcT2], f2 = oc[kOcF2];
  const double ta = oc[kOcTA], tb = oc[kOcTB];
  // still one-sided after several evaluations (a target the bracket may not
  // contain): the far end of the trial interval decides every remaining trial
  // if the root is not inside it
  if (n >= 6.0 && hasA != hasB) return hasA ? hi : lo;
  double t = 0.0;
  bool have = false;
  if (n >= 2.0 && f1 != f2) {
    t = t1 - f1 * (t1 - t2) / (f1 - f2);  // secant through the last two evaluations
    have = true;
  } else if (n == 1.0 && oc[kOcSlope] < 0.0) {
    t = t1 - f1 / oc[kOcSlope];  // Newton step with the previous bisection's slope
    have = true;
  }
  if (have && hasA && hasB) {
    const double tau = fmin(kOcTau, 0.25 * (tb - ta));
    // outside the bracket: the Illinois point of [A, B] instead
    if (!(t > ta && t < tb)) t = ta - oc[kOcFA] * (tb - ta) / (oc[kOcFB] - oc[kOcFA]);
    // an estimate next to the last evaluation: step just past it, across the root
    if (fabs(t - t1) < tau) t = t1 + (f1 > 0.0 ? tau : -tau);
    t = fmin(fmax(t, ta + tau), tb - tau);
  } else if (have) {
    if (hasA) t = fmax(t, ta);
    if (hasB) t = fmin(t, tb);
  }
  // a candidate log multiplier, as a multiplier strictly inside the trial
  // interval and the bracket (0 if it is not)
  auto inside = [&](double u) -> double {
    if (!isfinite(u) || (hasA && !(u > ta)) || (hasB && !(u < tb))) return 0.0;
    const double c = exp(u);
    return (c > lo && c < hi && (!hasA || c > oc[kOcA]) && (!hasB || c < oc[kOcB])) ? c : 0.0;
  };
  if (have) {
    const double c = inside(t);
    if (c > 0.0) return c;
  }
  if (n == 1.0 && hasA != hasB) {
    // one side known and no usable slope: a fixed log step toward the root
    const double c = inside(t1 + (hasA ? oc[kOcGrow] : -oc[kOcGrow]));
    if (c > 0.0) return c;
  }
  // one-sided and the secant leaves the interval: test its far end
  if (n >= 2.0 && hasA != hasB) return hasA ? hi : lo;
  return lam;
}

// One bisection step for the evaluation vol = vol(oc[kOcLam]) (one thread).
__device__ void oc_step(double* oc, double vol) {
  const double lamE = oc[kOcLam], f = vol - oc[kOcVt];
  oc[kOcVol] = vol;
  oc[kOcEvals] += 1.0;
  oc[kOcT2] = oc[kOcT1];
  oc[kOcF2] = oc[kOcF1];
  oc[kOcT1] = log(lamE);
  oc[kOcF1] = f;
  oc_record(oc, lamE, oc[kOcT1], f);
  double lo = oc[kOcLo], hi = oc[kOcHi];
  for (int guard = 0; guard < 200; ++guard) {
    const double lam = sqrt(lo * hi);  // the reference's next trial
    int dec = 0;                        // +1: vol(lam) > vt, -1: vol(lam) <= vt
    if (oc[kOcHasA] != 0.0 && lam <= oc[kOcA]) dec = 1;
    else if (oc[kOcHasB] != 0.0 && lam >= oc[kOcB]) dec = -1;
    if (dec != 0) {
      if (dec > 0) lo = lam;
      else hi = lam;
      oc[kOcLamLast] = lam;
      oc[kOcSteps] += 1.0;
      if (hi / lo < 1.0 + 1e-9 || oc[kOcSteps] >= 100.0) {
        oc[kOcDone] = 1.0;
        break;
      }
      continue;
    }
    oc[kOcLam] = oc_choose(oc, lo, hi, lam);
    break;
  }
  oc[kOcLo] = lo;
  oc[kOcHi] = hi;
  if (oc[kOcDone] != 0.0) {
    // the slope handed to the next bisection: from the last two evaluations
    const double dt = oc[kOcT1] - oc[kOcT2], s = dt != 0.0 ? (oc[kOcF1] - oc[kOcF2]) / dt : 0.0;
    oc[kOcSlopeNext] = (oc[kOcEvals] >= 2.0 && s < 0.0 && isfinite(s)) ? s : oc[kOcSlope];
  }
}

// number of (dy', dx') offsets in [-R, R]^2 with dy'^2 + dx'^2 = m
__host__ __device__ constexpr int ring_taps(int R, int m) {
  int n = 0;
  for (int dy = -R; dy <= R; ++dy)
    for (int dx = -R; dx <= R; ++dx) n += (dy * dy + dx * dx == m) ? 1 : 0;
  return n;
}

// kOc prologue of a filter launch: the step of the previous evaluation (the
// same bits in every workgroup, from the fixed-order sum of its partials),
// the state written out by workgroup 0, and this launch's trial parameters.
// Returns false once the bisection has stopped (the launch then only carries
// the state forward).  256 threads.
__device__ __forceinline__ bool oc_prologue(const FilterArgs& a, double* st, double* sred, double& rlam, double& mv) {
  const int tid = threadIdx.x;
  if (a.prev && a.oc[kOcDone] == 0.0) {
    const double vol = oc_volume(a.part_in, gridDim.x, sred);
    if (tid == 0) {
      for (int i = 0; i < kOcLen; ++i) st[i] = a.oc[i];
      oc_step(st, vol);
    }
  } else if (tid < kOcLen) {
    st[tid] = a.oc[tid];
  }
  __syncthreads();
  if (blockIdx.x == 0 && tid < kOcLen) a.oc_out[tid] = st[tid];
  if (st[kOcDone] != 0.0) return false;
  rlam = 1.0 / sqrt(st[kOcLam]);
  mv = a.beta <= 1.0 ? a.move : 0.5 * a.move;
  __syncthreads();  // sred is reused by the volume sum at the end
  return true;
}

template <int R, int MODE>
__global__ __launch_bounds__(FTX* FTY) void k_filter(FilterArgs a) {
  constexpr int K = 2 * R + 1, LX = FTX + 2 * R, LY = FTY + 2 * R, NP = LX * LY, NT = FTX * FTY;
  constexpr int NLD = (NP + NT - 1) / NT;  // staging loads per thread per layer
  constexpr int NS = FZC + 2 * R;          // input layers per chunk
  constexpr int C = MODE == kBwd2 ? 2 : 1; // channels filtered together
  __shared__ double buf[2][C][NP];
  __shared__ double sred[NT / 64];
  const int tid = threadIdx.x;
  double rlam = 0.0, mv = 0.0;
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const int ntx = (nx + FTX - 1) / FTX, nty = (ny + FTY - 1) / FTY;
  int bid = xcd_logical(blockIdx.x, gridDim.x);  // neighbouring tiles share one XCD's L2
  const int bx = bid % ntx;
  bid /= ntx;
  const int by = bid % nty;
  const int bz = bid / nty;
  const int X0 = bx * FTX, Y0 = by * FTY, Z0 = bz * FZC;
  const int64_t exy = (int64_t)nx * ny;

  // this thread's staging slots: in-plane element offset, or -1 outside the grid
  int soff[NLD];
#pragma unroll
  for (int k = 0; k < NLD; ++k) {
    const int idx = tid + k * NT;
    const int lx = idx % LX, ly = idx / LX;
    const int gx = X0 - R + lx, gy = Y0 - R + ly;
    soff[k] = (idx < NP && gx >= 0 && gx < nx && gy >= 0 && gy < ny) ? gx + nx * gy : -1;
  }
  // raw operands of one staged value (flags, input, second input)
  struct Raw {
    uint8_t f;
    double x, y, x2;
  };
  Raw pre[2][NLD];
  auto fetch = [&](int q, int slot) {
    const bool qin = q >= 0 && q < nz;
#pragma unroll
    for (int k = 0; k < NLD; ++k) {
      Raw& v = pre[slot][k];
      v.f = 0;
      v.x = 0.0;
      v.y = 0.0;
      v.x2 = 0.0;
      if (qin && soff[k] >= 0) {
        const int64_t e = soff[k] + exy * q;
        v.f = a.flags[e];
        if (MODE != kHs) v.x = a.in[e];
        if (MODE == kBwd || MODE == kBwd2) v.y = a.invHs[e];
        if (MODE == kBwd2) v.x2 = a.in2[e];
        if (MODE == kOc) v.y = a.dcx[e];
      }
    }
  };
  auto staged = [&](const Raw& v, int ch) -> double {
    if (!(v.f & kActive)) return 0.0;
    if (MODE == kHs) return 1.0;
    if (MODE == kFwd || MODE == kPhys) return v.x;
    if (MODE == kBwd || MODE == kBwd2) return (ch == 0 ? v.x : v.x2) * v.y;
    // kOc: trial design of optimize.py:150-153, with sqrt(c/dv) precomputed
    if (v.f & kFrozen) return 1.0;
    return fmin(fmax(v.x * (v.y * rlam), fmax(0.0, v.x - mv)), fmin(1.0, v.x + mv));
  };

  const int tx = tid % FTX, ty = tid / FTX;
  const int gx = X0 + tx, gy = Y0 + ty;
  const bool col = gx < nx && gy < ny;
  const int64_t coff = gx + (int64_t)nx * gy;
  struct Epi {
    uint8_t f;
    double ih, ph;
  };
  auto fetch_epi = [&](int j, Epi& E) {
    const int z = Z0 + j;
    E.f = 0;
    E.ih = 0.0;
    E.ph = 0.0;
    if (col && j < FZC && z < nz) {
      const int64_t e = coff + exy * z;
      E.f = a.flags[e];
      if (MODE == kFwd || MODE == kPhys || MODE == kOc) E.ih = a.invHs[e];
      if (MODE == kPhys || MODE == kOc) E.ph = a.phi[e];
    }
  };

  // win[c][i] accumulates output layer s - 2R + i of channel c while input layer s is processed
  double win[C][K];
#pragma unroll
  for (int c = 0; c < C; ++c)
#pragma unroll
    for (int i = 0; i < K; ++i) win[c][i] = 0.0;
  double vol = 0.0;
  double dst[3] = {0.0, 0.0, 0.0};  // kBwd2: design count, sum of -out, sum of out2
  fetch(Z0 - R, 0);
  fetch(Z0 - R + 1, 1);
  Epi ep[2] = {};
  if (MODE == kOc) {
    // the bisection step of the previous evaluation (identical in every
    // workgroup; the first two layers' loads are in flight meanwhile), then
    // this launch's trial multiplier
    static_assert(NT == 256, "oc_volume sums over 256 threads");
    __shared__ double st[kOcLen];
    if (!oc_prologue(a, st, sred, rlam, mv)) return;
  }

  // one input layer; SL = s & 1 selects the LDS buffer and the prefetch slots
  auto step = [&](int s, auto slc) {
    constexpr int SL = decltype(slc)::value;
#pragma unroll
    for (int k = 0; k < NLD; ++k) {
      const int idx = tid + k * NT;
      if (idx < NP) {
#pragma unroll
        for (int c = 0; c < C; ++c) buf[SL][c][idx] = staged(pre[SL][k], c);
      }
    }
    if (s + 2 < NS) fetch(Z0 - R + s + 2, SL);
    // output layer jo = s - 2R is finished in this step (its operands were fetched
    // one step earlier, in slot jo & 1 = SL); fetch the next layer's
    const int jo = s - 2 * R;
    const Epi E = ep[SL];
    if (jo + 1 >= 0 && jo + 1 < FZC) fetch_epi(jo + 1, ep[SL ^ 1]);
    __syncthreads();
#pragma unroll
    for (int c = 0; c < C; ++c) {
      const double* b = buf[SL][c];
      // ring sums of the (dy, dx) window by squared in-plane distance m = dy'^2 + dx'^2
      double P[2 * R * R + 1];
#pragma unroll
      for (int m = 0; m <= 2 * R * R; ++m) P[m] = 0.0;
#pragma unroll
      for (int dy = 0; dy < K; ++dy)
#pragma unroll
        for (int dx = 0; dx < K; ++dx) P[(dy - R) * (dy - R) + (dx - R) * (dx - R)] += b[(ty + dy) * LX + tx + dx];
      // the layer reaches output layers s - R +- o (window slots R +- o) with weight
      // w(o^2 + m) on ring m
#pragma unroll
      for (int o = 0; o <= R; ++o) {
        double q = 0.0;
#pragma unroll
        for (int m = 0; m <= 2 * R * R; ++m)
          if (ring_taps(R, m) > 0) q += a.wc[o * o + m] * P[m];
        win[c][R + o] += q;
        if (o > 0) win[c][R - o] += q;
      }
    }
    const int z = Z0 + jo;
    if (jo >= 0 && col && z < nz) {
      const int64_t e = coff + exy * z;
      const bool act = E.f & kActive;
      const double v = win[0][0];
      if (MODE == kHs) {
        a.out[e] = v;
      } else if (MODE == kFwd) {
        a.out[e] = act ? v * E.ih : 0.0;
      } else if (MODE == kBwd) {
        a.out[e] = act ? v : 0.0;
      } else if (MODE == kBwd2) {
        const double v2 = win[C - 1][0];
        a.out[e] = act ? v : 0.0;
        a.out2[e] = act ? v2 : 0.0;
        if (act && !(E.f & kFrozen)) {
          dst[0] += 1.0;
          dst[1] += -v;
          dst[2] += v2;
        }
      } else {
        double xp = 0.0, dxp = 0.0;
        if (act) {
          project(v * E.ih, a, xp, dxp);
          if (E.f & kFrozen) xp = 1.0;
          vol += E.ph * xp;
        }
        if (MODE == kPhys) {
          a.out[e] = xp;
          a.out2[e] = dxp;
        }
      }
    }
#pragma unroll
    for (int c = 0; c < C; ++c) {
#pragma unroll
      for (int i = 0; i + 1 < K; ++i) win[c][i] = win[c][i + 1];
      win[c][K - 1] = 0.0;
    }
  };
  static_assert(NS % 2 == 0, "the step loop handles input layers in pairs");
#pragma unroll 1
  for (int s = 0; s < NS; s += 2) {
    step(s, std::integral_constant<int, 0>{});
    step(s + 1, std::integral_constant<int, 1>{});
  }
  if (MODE == kPhys || MODE == kOc) {
    for (int off = 32; off > 0; off >>= 1) vol += __shfl_down(vol, off, 64);
    if ((tid & 63) == 0) sred[tid >> 6] = vol;
    __syncthreads();
    if (tid == 0) {
      double sm = 0.0;
      for (int w = 0; w < NT / 64; ++w) sm += sred[w];
      a.part[blockIdx.x] = sm;
    }
  }
  if (MODE == kBwd2 && a.part) {
    // partials in the layout of the design statistics: [q][block], q = 3 (count), 4, 5; 0-2 unused
    __shared__ double sst[3][NT / 64];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      double v = dst[q];
      for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
      if ((tid & 63) == 0) sst[q][tid >> 6] = v;
    }
    __syncthreads();
    if (tid < 6) {
      double sm = 0.0;
      if (tid >= 3)
        for (int w = 0; w < NT / 64; ++w) sm += sst[tid - 3][w];
      a.part[tid * gridDim.x + blockIdx.x] = sm;
    }
  }
}

// Gather path for radii above kRingMaxR, first kernel: the staged value of
// every element (what k_filter writes to LDS) into a.stg, grid-stride over
// the same number of workgroups as the gather kernel, so that an OC
// evaluation's prologue sums as many volume partials as the previous
// evaluation's gather wrote.
template <int MODE>
__global__ __launch_bounds__(256) void k_filter_stage(FilterArgs a) {
  __shared__ double st[kOcLen];
  __shared__ double sred[4];
  double rlam = 0.0, mv = 0.0;
  if (MODE == kOc && !oc_prologue(a, st, sred, rlam, mv)) return;
  const int64_t n = (int64_t)a.nx * a.ny * a.nz;
  for (int64_t e = blockIdx.x * 256 + threadIdx.x; e < n; e += (int64_t)gridDim.x * 256) {
    const uint8_t f = a.flags[e];
    double v = 0.0;
    if (f & kActive) {
      if (MODE == kHs) v = 1.0;
      else if (MODE == kFwd || MODE == kPhys) v = a.in[e];
      else if (MODE == kBwd) v = a.in[e] * a.invHs[e];
      else if (f & kFrozen) v = 1.0;
      else {
        const double x = a.in[e];
        v = fmin(fmax(x * (a.dcx[e] * rlam), fmax(0.0, x - mv)), fmin(1.0, x + mv));
      }
    }
    a.stg[e] = v;
  }
}

// Gather path, second kernel: each output sums w(dz^2 + dy^2 + dx^2) times
// the staged values of its (2R+1)^3 cube (inside the grid) in a fixed order
// (dz, dy, dx ascending), then the epilogue of k_filter.  A workgroup covers
// an FTX x FTY x FZC tile (one column per thread), as k_filter does.
template <int MODE>
__global__ __launch_bounds__(FTX* FTY) void k_filter_gather(FilterArgs a) {
  constexpr int NT = FTX * FTY;
  __shared__ double sred[NT / 64];
  if (MODE == kOc && a.oc_out[kOcDone] != 0.0) return;  // stopped: the stage launch carried the state
  const int tid = threadIdx.x, R = a.R;
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const int ntx = (nx + FTX - 1) / FTX, nty = (ny + FTY - 1) / FTY;
  int bid = xcd_logical(blockIdx.x, gridDim.x);
  const int bx = bid % ntx;
  bid /= ntx;
  const int by = bid % nty;
  const int bz = bid / nty;
  const int gx = bx * FTX + tid % FTX, gy = by * FTY + tid / FTX;
  const bool col = gx < nx && gy < ny;
  const int64_t exy = (int64_t)nx * ny;
  const int x0 = max(gx - R, 0), x1 = min(gx + R, nx - 1), y0 = max(gy - R, 0), y1 = min(gy + R, ny - 1);
  double vol = 0.0;
  for (int j = 0; j < FZC; ++j) {
    const int z = bz * FZC + j;
    if (!col || z >= nz) break;
    double v = 0.0;
    for (int zz = max(z - R, 0); zz <= min(z + R, nz - 1); ++zz) {
      const int dz2 = (zz - z) * (zz - z);
      for (int yy = y0; yy <= y1; ++yy) {
        const int d2 = dz2 + (yy - gy) * (yy - gy);
        const double* row = a.stg + exy * zz + (int64_t)nx * yy;
        for (int xx = x0; xx <= x1; ++xx) v += a.wt[d2 + (xx - gx) * (xx - gx)] * row[xx];
      }
    }
    const int64_t e = gx + (int64_t)nx * gy + exy * z;
    const uint8_t f = a.flags[e];
    const bool act = f & kActive;
    if (MODE == kHs) {
      a.out[e] = v;
    } else if (MODE == kFwd) {
      a.out[e] = act ? v * a.invHs[e] : 0.0;
    } else if (MODE == kBwd) {
      a.out[e] = act ? v : 0.0;
    } else {
      double xp = 0.0, dxp = 0.0;
      if (act) {
        project(v * a.invHs[e], a, xp, dxp);
        if (f & kFrozen) xp = 1.0;
        vol += a.phi[e] * xp;
      }
      if (MODE == kPhys) {
        a.out[e] = xp;
        a.out2[e] = dxp;
      }
    }
  }
  if (MODE == kPhys || MODE == kOc) {
    for (int off = 32; off > 0; off >>= 1) vol += __shfl_down(vol, off, 64);
    if ((tid & 63) == 0) sred[tid >> 6] = vol;
    __syncthreads();
    if (tid == 0) {
      double sm = 0.0;
      for (int w = 0; w < NT / 64; ++w) sm += sred[w];
      a.part[blockIdx.x] = sm;
    }
  }
}

// The last evaluation of a launch batch: its step, in place (one workgroup).
// The state also goes to out_host (device alias of pinned host memory).
__global__ __launch_bounds__(256) void k_oc_bisect(const double* __restrict__ part, int nblk, double* __restrict__ oc,
                                                   double* __restrict__ out_host) {
  __shared__ double sh[4];
  if (oc[kOcDone] == 0.0) {
    const double vol = oc_volume(part, nblk, sh);
    if (threadIdx.x == 0) oc_step(oc, vol);
    __syncthreads();
  }
  if (threadIdx.x < kOcLen) out_host[threadIdx.x] = oc[threadIdx.x];
}

// sqrt(max(-dcx, 0) / dvx) per element, so a trial needs one multiply by
// lam^-1/2 (dvx = 0 with -dcx > 0 gives +inf, which the move limit clips).
// Block 0 also brings the initial bisection state in from pinned host memory.
// With stats (the design statistics of design_stats), block 0 also sets the
// bracket and the first multiplier from them: the reference's bracket scale
// (optimize.py:146-147) with lo = 1e-8 scale, hi = 1e8 scale, and the first
// evaluation at scale * ratio when that lies inside (else at sqrt(lo hi)).
__global__ void k_oc_prep(const double* __restrict__ dcx, const double* __restrict__ dvx, int64_t n,
                          double* __restrict__ r0, const double* __restrict__ st_host, double* __restrict__ st,
                          const double* __restrict__ stats, double ratio) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    double v[kOcLen];
    for (int i = 0; i < kOcLen; ++i) v[i] = st_host[i];
    if (stats) {
      const double nd = stats[3];
      const double scale = nd > 0.0 ? fmax((stats[4] / nd) / fmax(stats[5] / nd, 1e-30), 1e-300) : 1.0;
      const double lo = scale * 1e-8, hi = scale * 1e8, hint = scale * ratio;
      v[kOcLo] = lo;
      v[kOcHi] = hi;
      v[kOcLam] = (ratio > 0.0 && hint > lo && hint < hi) ? hint : sqrt(lo * hi);
    }
    for (int i = 0; i < kOcLen; ++i) st[i] = v[i];
  }
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= n) return;
  const double c = fmax(-dcx[e], 0.0);
  r0[e] = c > 0.0 ? sqrt(c / dvx[e]) : 0.0;
}

__global__ void k_oc_final(const double* __restrict__ x, const double* __restrict__ dcx, const double* __restrict__ dvx,
                           const uint8_t* __restrict__ flags, const double* __restrict__ oc, double beta, double move,
                           int64_t n, double* __restrict__ xnew) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= n) return;
  const uint8_t f = flags[e];
  double v = 0.0;
  if (f & kActive) {
    if (f & kFrozen) {
      v = 1.0;
    } else {
      const double lam = oc[kOcLamLast];
      const double mv = beta <= 1.0 ? move : 0.5 * move;
      const double xv = x[e];
      const double ratio = sqrt(fmax(-dcx[e], 0.0) / fmax(dvx[e] * lam, 1e-300));
      v = fmin(fmax(xv * ratio, fmax(0.0, xv - mv)), fmin(1.0, xv + mv));
    }
  }
  xnew[e] = v;
}

__global__ void k_inv_hs(double* __restrict__ Hs, double* __restrict__ invHs, int64_t n) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= n) return;
  double h = Hs[e];
  if (h == 0.0) h = 1.0;
  Hs[e] = h;
  invHs[e] = 1.0 / h;
}

// ---------------------------------------------------------------------------
// hat weight per squared voxel distance d2 = 0 .. 3R^2 (optimize.py:28-31: k = max(0, r - |d|))
static std::vector<double> cone_weights(double r, int R) {
  std::vector<double> w(3 * R * R + 1);
  for (int d2 = 0; d2 < (int)w.size(); ++d2) w[d2] = std::max(0.0, r - std::sqrt(double(d2)));
  return w;
}

template <int MODE>
static void launch_filter_mode(int R, const FilterArgs& a, int nblk, hipStream_t s) {
  switch (R) {
    case 1: hipLaunchKernelGGL((k_filter<1, MODE>), dim3(nblk), dim3(FTX * FTY), 0, s, a); break;
    case 2: hipLaunchKernelGGL((k_filter<2, MODE>), dim3(nblk), dim3(FTX * FTY), 0, s, a); break;
    case 3: hipLaunchKernelGGL((k_filter<3, MODE>), dim3(nblk), dim3(FTX * FTY), 0, s, a); break;
    case 4: hipLaunchKernelGGL((k_filter<4, MODE>), dim3(nblk), dim3(FTX * FTY), 0, s, a); break;
    case 5: hipLaunchKernelGGL((k_filter<5, MODE>), dim3(nblk), dim3(FTX * FTY), 0, s, a); break;
    default:
      static_assert(kRingMaxR == 5, "one case per ring-kernel radius");
      hipLaunchKernelGGL((k_filter_stage<MODE>), dim3(nblk), dim3(256), 0, s, a);
      ANNEAL_HIP(hipGetLastError());
      hipLaunchKernelGGL((k_filter_gather<MODE>), dim3(nblk), dim3(FTX * FTY), 0, s, a);
      break;
  }
  ANNEAL_HIP(hipGetLastError());
}
static void launch_filter(int mode, int R, const FilterArgs& a, int nblk, hipStream_t s) {
  switch (mode) {
    case kHs: launch_filter_mode<kHs>(R, a, nblk, s); break;
    case kFwd: launch_filter_mode<kFwd>(R, a, nblk, s); break;
    case kBwd: launch_filter_mode<kBwd>(R, a, nblk, s); break;
    case kPhys: launch_filter_mode<kPhys>(R, a, nblk, s); break;
    case kOc: launch_filter_mode<kOc>(R, a, nblk, s); break;
  }
}

FilterContext::FilterContext(int nx, int ny, int nz, double r_vox, const uint8_t* flags_host, hipStream_t stream)
    : g_{nx, ny, nz}, s_(stream) {
  ANNEAL_CHECK(nx >= 1 && ny >= 1 && nz >= 1, "grid dimensions must be positive");
  ANNEAL_CHECK(r_vox > 0.0 && std::isfinite(r_vox), "filter radius must be positive");
  R_ = std::max((int)std::ceil(r_vox) - 1, 1);
  w_ = cone_weights(r_vox, R_);
  const int64_t n = g_.nelem();
  nblk_ = ((nx + FTX - 1) / FTX) * ((ny + FTY - 1) / FTY) * ((nz + FZC - 1) / FZC);
  ANNEAL_HIP(hipMalloc(&flags_, n));
  ANNEAL_HIP(hipMalloc(&Hs_, n * sizeof(double)));
  ANNEAL_HIP(hipMalloc(&invHs_, n * sizeof(double)));
  ANNEAL_HIP(hipMalloc(&r0_, n * sizeof(double)));
  ANNEAL_HIP(hipMalloc(&part_, std::max(6 * nblk_, 6 * 512) * sizeof(double)));
  ANNEAL_HIP(hipMalloc(&oc_, 2 * kOcLen * sizeof(double)));
  ANNEAL_HIP(hipHostMalloc((void**)&host_, 2 * kOcLen * sizeof(double), hipHostMallocDefault));
  ANNEAL_HIP(hipHostGetDevicePointer((void**)&host_dev_, host_, 0));
  ANNEAL_HIP(hipEventCreateWithFlags(&oc_ready_, hipEventDisableTiming));
  ANNEAL_HIP(hipMalloc(&wt_, w_.size() * sizeof(double)));
  ANNEAL_HIP(hipMemcpy(wt_, w_.data(), w_.size() * sizeof(double), hipMemcpyHostToDevice));
  if (R_ > kRingMaxR) ANNEAL_HIP(hipMalloc(&stg_, n * sizeof(double)));
  ANNEAL_HIP(hipMemcpy(flags_, flags_host, n, hipMemcpyHostToDevice));
  FilterArgs a = args();
  a.out = Hs_;
  launch_filter(kHs, R_, a, nblk_, s_);
  hipLaunchKernelGGL(k_inv_hs, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s_, Hs_, invHs_, n);
  ANNEAL_HIP(hipGetLastError());
  ANNEAL_HIP(hipStreamSynchronize(s_));
}

FilterContext::~FilterContext() {
  (void)hipStreamSynchronize(s_);
  (void)hipFree(flags_);
  (void)hipFree(Hs_);
  (void)hipFree(invHs_);
  (void)hipFree(r0_);
  (void)hipFree(part_);
  (void)hipFree(oc_);
  (void)hipFree(wt_);
  (void)hipFree(stg_);
  (void)hipHostFree(host_);
  (void)hipEventDestroy(oc_ready_);
}

FilterArgs FilterContext::args() const {
  FilterArgs a{};
  a.flags = flags_;
  a.invHs = invHs_;
  std::copy(w_.begin(), w_.begin() + std::min<size_t>(w_.size(), kWc), a.wc);
  a.wt = wt_;
  a.stg = stg_;
  a.R = R_;
  a.nx = g_.nx;
  a.ny = g_.ny;
  a.nz = g_.nz;
  return a;
}

void FilterContext::forward(const double* x, double* y) {
  FilterArgs a = args();
  a.in = x;
  a.out = y;
  launch_filter(kFwd, R_, a, nblk_, s_);
}

void FilterContext::backward(const double* d, double* y) {
  FilterArgs a = args();
  a.in = d;
  a.out = y;
  launch_filter(kBwd, R_, a, nblk_, s_);
}

void FilterContext::physical(const double* x, double beta, const double* phi, double* xphys, double* dxp,
                             double* vol_dev) {
  FilterArgs a = args();
  a.in = x;
  a.phi = phi;
  a.out = xphys;
  a.out2 = dxp;
  a.part = part_;
  set_beta(a, beta);
  launch_filter(kPhys, R_, a, nblk_, s_);
  if (vol_dev) finalize(part_, nblk_, 1, vol_dev, s_);
}

void FilterContext::oc_begin(const double* x, const double* dcx, const double* dvx, const double* phi, double beta,
                             double move, double lo, double hi, double vtarget, double* xnew, double hint,
                             double width) {
  oc_start(x, dcx, dvx, phi, beta, move, lo, hi, vtarget, xnew, hint, width, nullptr, 0.0);
}

void FilterContext::oc_begin_stats(const double* x, const double* dcx, const double* dvx, const double* phi,
                                   double beta, double move, const double* stats, double ratio, double vtarget,
                                   double* xnew, double width) {
  oc_start(x, dcx, dvx, phi, beta, move, 1.0, 1.0, vtarget, xnew, 0.0, width, stats, ratio);
}

void FilterContext::oc_start(const double* x, const double* dcx, const double* dvx, const double* phi, double beta,
                             double move, double lo, double hi, double vtarget, double* xnew, double hint,
                             double width, const double* stats, double ratio) {
  if (run_.active) {
    // an abandoned bisection (its caller stopped before oc_end): let its
    // launches finish before the pinned state is rewritten
    ANNEAL_HIP(hipStreamSynchronize(s_));
    run_.active = false;
  }
  // the initial state goes up from the pinned upper half of host_ (k_oc_prep
  // reads it, possibly after this call returns)
  double* st = host_ + kOcLen;
  std::fill(st, st + kOcLen, 0.0);
  st[kOcLo] = lo;
  st[kOcHi] = hi;
  st[kOcLam] = (hint > lo && hint < hi) ? hint : std::sqrt(lo * hi);
  st[kOcVt] = vtarget;
  st[kOcGrow] = width > 0.0 ? std::min(width, 0.5) : kOcGrowStep;
  st[kOcSlope] = oc_slope_;
  const int64_t ne = g_.nelem();
  hipLaunchKernelGGL(k_oc_prep, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, s_, dcx, dvx, ne, r0_,
                     host_dev_ + kOcLen, oc_, stats, ratio);
  ANNEAL_HIP(hipGetLastError());
  run_ = OcRun{x, phi, xnew, dcx, dvx, beta, move, 0, true, oc_};
  oc_launch(kOcFirstBatch);
}

void FilterContext::oc_launch(int count) {
  FilterArgs a = args();
  a.in = run_.x;
  a.dcx = r0_;
  a.dvx = nullptr;
  a.phi = run_.phi;
  set_beta(a, run_.beta);
  a.move = run_.move;
  // Evaluation e reads the state in oc_[e % 2], first takes the step of
  // evaluation e - 1 from the partials in part_[(e - 1) % 2] (except the first
  // of a batch, whose predecessor the batch finisher took), evaluates and
  // writes its partials to part_[e % 2] and the state to oc_[(e + 1) % 2].
  // Evaluations launched after the stop only carry the state forward.
  int batch = 0;
  for (; batch < count && run_.launched < 240; ++batch, ++run_.launched) {
    const int e = run_.launched;
    a.oc = oc_ + (e % 2) * kOcLen;
    a.oc_out = oc_ + ((e + 1) % 2) * kOcLen;
    a.part = part_ + (e % 2) * nblk_;
    a.part_in = part_ + ((e + 1) % 2) * nblk_;
    a.prev = batch > 0;
    launch_filter(kOc, R_, a, nblk_, s_);
  }
  if (batch == 0) return;
  // the batch finisher: the last evaluation's step, in place
  const int e = run_.launched;
  run_.state = oc_ + (e % 2) * kOcLen;
  hipLaunchKernelGGL(k_oc_bisect, dim3(1), dim3(256), 0, s_, part_ + ((e + 1) % 2) * nblk_, nblk_, run_.state,
                     host_dev_);
  ANNEAL_HIP(hipGetLastError());
  ANNEAL_HIP(hipEventRecord(oc_ready_, s_));
}

bool FilterContext::oc_poll() {
  if (!run_.active) throw std::runtime_error("oc_poll without oc_begin");
  ANNEAL_HIP(hipEventSynchronize(oc_ready_));
  if (host_[kOcDone] != 0.0) return true;
  if (run_.launched >= 240) {
    run_.active = false;
    throw std::runtime_error("OC bisection did not terminate");
  }
  oc_launch(kOcNextBatch);
  return false;
}

int FilterContext::oc_end(int* evals, double* lam_last) {
  if (!run_.active || host_[kOcDone] == 0.0) throw std::runtime_error("oc_end before the bisection has stopped");
  run_.active = false;
  const int64_t n = g_.nelem();
  hipLaunchKernelGGL(k_oc_final, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s_, run_.x, run_.dcx, run_.dvx,
                     flags_, run_.state, run_.beta, run_.move, n, run_.xnew);
  ANNEAL_HIP(hipGetLastError());
  oc_slope_ = host_[kOcSlopeNext];
  if (evals) *evals = (int)host_[kOcEvals];
  if (lam_last) *lam_last = host_[kOcLamLast];
  return (int)host_[kOcSteps];
}

int FilterContext::oc_bisect(const double* x, const double* dcx, const double* dvx, const double* phi, double beta,
                             double move, double lo, double hi, double vtarget, double* xnew, double hint,
                             double width, int* evals, double* lam_last) {
  oc_begin(x, dcx, dvx, phi, beta, move, lo, hi, vtarget, xnew, hint, width);
  while (!oc_poll()) {
  }
  return oc_end(evals, lam_last);
}

// ---------------------------------------------------------------------------
// Element energies in the Walsh basis: ce = sum over modes of u_hat * (Khat u_hat).
__device__ __forceinline__ void element_modes(const double* __restrict__ u, int64_t ldn, int i, int j, int k, int nx,
                                              int ny, double (&S)[3][4], double (&D)[3][4]) {
  const int64_t sx = nx + 1, sxy = (int64_t)(nx + 1) * (ny + 1);
  const int64_t n0 = i + sx * j + sxy * k;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const double* p = u + c * ldn;
    double Fb[4], Ft[4];
    face_wht(p[n0], p[n0 + 1], p[n0 + sx], p[n0 + sx + 1], Fb);
    face_wht(p[n0 + sxy], p[n0 + sxy + 1], p[n0 + sxy + sx], p[n0 + sxy + sx + 1], Ft);
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      S[c][m] = Ft[m] + Fb[m];
      D[c][m] = Ft[m] - Fb[m];
    }
  }
}

__device__ __forceinline__ double element_ce(const double (&S)[3][4], const double (&D)[3][4],
                                             const KeCoef<double>& k) {
  double YS[3][4], YD[3][4];
  element_walsh_modes(S, D, k, YS, YD);
  double ce = 0.0;
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int m = 0; m < 4; ++m) ce += S[c][m] * YS[c][m] + D[c][m] * YD[c][m];
  return ce;
}

__global__ void k_element_energy(const double* __restrict__ U, int nrhs, int64_t ldn, int nx, int ny, int nz,
                                 KeCoef<double> k, const uint8_t* __restrict__ flags, double* __restrict__ out) {
  const int64_t ne = (int64_t)nx * ny * nz;
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= ne) return;
  const int i = e % nx, j = (e / nx) % ny, kk = e / ((int64_t)nx * ny);
  const bool act = flags[e] & kActive;
  for (int r = 0; r < nrhs; ++r) {
    double ce = 0.0;
    if (act) {
      double S[3][4], D[3][4];
      element_modes(U + (int64_t)3 * r * ldn, ldn, i, j, kk, nx, ny, S, D);
      ce = element_ce(S, D, k);
    }
    out[(int64_t)r * ne + e] = ce;
  }
}

__global__ void k_sensitivity(const double* __restrict__ U, int nrhs, int64_t ldn, int nx, int ny, int nz,
                              KeCoef<double> k, const uint8_t* __restrict__ flags, const double* __restrict__ xphys,
                              const double* __restrict__ dxp, const double* __restrict__ phi,
                              const double* __restrict__ coef, double p, double emin, double E0,
                              double* __restrict__ dcdx, double* __restrict__ dvdx) {
  const int64_t ne = (int64_t)nx * ny * nz;
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= ne) return;
  if (!(flags[e] & kActive)) {
    dcdx[e] = 0.0;
    dvdx[e] = 0.0;
    return;
  }
  const int i = e % nx, j = (e / nx) % ny, kk = e / ((int64_t)nx * ny);
  const double fac = -p * pow(xphys[e], p - 1.0) * (1.0 - emin) * E0 * phi[e];
  double dc = 0.0;
  for (int r = 0; r < nrhs; ++r) {
    double S[3][4], D[3][4];
    element_modes(U + (int64_t)3 * r * ldn, ldn, i, j, kk, nx, ny, S, D);
    dc += coef[r] * (fac * element_ce(S, D, k));
  }
  dc = fmin(dc, 0.0);
  dcdx[e] = dc * dxp[e];
  dvdx[e] = phi[e] * dxp[e];
}

__global__ void k_stiffness_law(const double* __restrict__ xphys, const double* __restrict__ phi,
                                const uint8_t* __restrict__ flags, double E0, double emin, double p, int64_t n,
                                double* __restrict__ Ee) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= n) return;
  Ee[e] = (flags[e] & kActive) ? E0 * phi[e] * (emin + pow(xphys[e], p) * (1.0 - emin)) : 0.0;
}

constexpr int kStatBlocks = 512;
__global__ __launch_bounds__(256) void k_design_stats(const double* __restrict__ xold, const double* __restrict__ xnew,
                                                      const double* __restrict__ xphys, const double* __restrict__ phi,
                                                      const double* __restrict__ dcx, const double* __restrict__ dvx,
                                                      const uint8_t* __restrict__ flags, int64_t n,
                                                      double* __restrict__ part) {
  double s[6] = {0, 0, 0, 0, 0, 0};
  // four grid-stride elements per pass, all loads issued first (clamped
  // addresses); the accumulation order is the element order of a plain loop
  constexpr int U = 4;
  const int64_t stride = (int64_t)gridDim.x * 256;
  for (int64_t e0 = blockIdx.x * 256 + threadIdx.x; e0 < n; e0 += U * stride) {
    uint8_t f[U];
    double vp[U], vx[U], vo[U], vn[U], vc[U], vv[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t e = e0 + u * stride;
      const int64_t c = e < n ? e : n - 1;
      f[u] = e < n ? flags[c] : uint8_t(0);
      vp[u] = xphys ? phi[c] : 0.0;
      vx[u] = xphys ? xphys[c] : 0.0;
      vo[u] = (xold && xnew) ? xold[c] : 0.0;
      vn[u] = (xold && xnew) ? xnew[c] : 0.0;
      vc[u] = dcx ? dcx[c] : 0.0;
      vv[u] = dvx ? dvx[c] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (!(f[u] & kActive)) continue;
      if (xphys) s[0] += vp[u] * vx[u];
      if (f[u] & kFrozen) continue;
      s[3] += 1.0;
      if (xphys && vx[u] > 0.1 && vx[u] < 0.9) s[1] += 1.0;
      if (xold && xnew) s[2] = fmax(s[2], fabs(vn[u] - vo[u]));
      if (dcx) s[4] += -vc[u];
      if (dvx) s[5] += vv[u];
    }
  }
  __shared__ double sh[6][4];
  for (int q = 0; q < 6; ++q) {
    double v = s[q];
    for (int off = 32; off > 0; off >>= 1) {
      const double o = __shfl_down(v, off, 64);
      v = (q == 2) ? fmax(v, o) : v + o;
    }
    if ((threadIdx.x & 63) == 0) sh[q][threadIdx.x >> 6] = v;
  }
  __syncthreads();
  if (threadIdx.x < 6) {
    const int q = threadIdx.x;
    double v = sh[q][0];
    for (int w = 1; w < 4; ++w) v = (q == 2) ? fmax(v, sh[q][w]) : v + sh[q][w];
    part[q * gridDim.x + blockIdx.x] = v;
  }
}
// One wavefront per statistic, fixed-order tree over the block partials.
__global__ __launch_bounds__(384) void k_design_stats_final(const double* __restrict__ part, int nblk,
                                                            double* __restrict__ out) {
  const int q = threadIdx.x >> 6, lane = threadIdx.x & 63;
  double v = 0.0;
  for (int b = lane; b < nblk; b += 64) v = (q == 2) ? fmax(v, part[q * nblk + b]) : v + part[q * nblk + b];
  for (int off = 32; off > 0; off >>= 1) {
    const double o = __shfl_down(v, off, 64);
    v = (q == 2) ? fmax(v, o) : v + o;
  }
  if (lane == 0) out[q] = v;
}

// Element-centre stress and von Mises (anneal/fem.py:244-263): eps = B0 u_e
// with the 24 element dofs in the reference's corner order (corner q at
// offset (a, b, c) = local[q]), sig = E D eps, Voigt xx yy zz xy yz zx.
__global__ void k_element_stress(const double* __restrict__ U, int nrhs, int64_t ldn, int nx, int ny, int nz,
                                 StressMats m, const uint8_t* __restrict__ flags, const double* __restrict__ E,
                                 double Es, double* __restrict__ sig, double* __restrict__ vm) {
  const int64_t ne = (int64_t)nx * ny * nz;
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= ne) return;
  const int i = e % nx, j = (e / nx) % ny, kk = e / ((int64_t)nx * ny);
  const bool act = flags[e] & kActive;
  const int64_t sx = nx + 1, sxy = (int64_t)(nx + 1) * (ny + 1);
  const double Ee = E ? E[e] : Es;
  for (int r = 0; r < nrhs; ++r) {
    double s[6] = {0, 0, 0, 0, 0, 0};
    if (act) {
      double ue[24];
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int64_t node = (i + m.local[q][0]) + sx * (j + m.local[q][1]) + sxy * (kk + m.local[q][2]);
#pragma unroll
        for (int c = 0; c < 3; ++c) ue[3 * q + c] = U[(int64_t)(3 * r + c) * ldn + node];
      }
      double eps[6];
#pragma unroll
      for (int a = 0; a < 6; ++a) {
        double t = 0.0;
#pragma unroll
        for (int b = 0; b < 24; ++b) t += m.B0[a][b] * ue[b];
        eps[a] = t;
      }
#pragma unroll
      for (int a = 0; a < 6; ++a) {
        double t = 0.0;
#pragma unroll
        for (int b = 0; b < 6; ++b) t += m.D[a][b] * eps[b];
        s[a] = t * Ee;
      }
    }
    double* o = sig + ((int64_t)r * ne + e) * 6;
#pragma unroll
    for (int a = 0; a < 6; ++a) o[a] = s[a];
    if (vm) {
      const double d0 = s[0] - s[1], d1 = s[1] - s[2], d2 = s[2] - s[0];
      vm[(int64_t)r * ne + e] =
          sqrt(0.5 * (d0 * d0 + d1 * d1 + d2 * d2) + 3.0 * (s[3] * s[3] + s[4] * s[4] + s[5] * s[5]));
    }
  }
}

void element_stress(const double* U, int nrhs, int64_t ldn, const GridDims& g, const StressMats& m,
                    const uint8_t* flags, const double* E, double Es, double* sig, double* vm, hipStream_t s) {
  const int64_t n = g.nelem();
  hipLaunchKernelGGL(k_element_stress, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, U, nrhs, ldn, g.nx, g.ny,
                     g.nz, m, flags, E, Es, sig, vm);
  ANNEAL_HIP(hipGetLastError());
}

void element_energy(const double* U, int nrhs, int64_t ldn, const GridDims& g, const KeCoef<double>& k,
                    const uint8_t* flags, double* out, hipStream_t s) {
  const int64_t n = g.nelem();
  hipLaunchKernelGGL(k_element_energy, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, U, nrhs, ldn, g.nx, g.ny,
                     g.nz, k, flags, out);
  ANNEAL_HIP(hipGetLastError());
}

void sensitivity(const double* U, int nrhs, int64_t ldn, const GridDims& g, const KeCoef<double>& k,
                 const uint8_t* flags, const double* xphys, const double* dxp, const double* phi, const double* coef,
                 double p, double emin, double E0, double* dcdx, double* dvdx, hipStream_t s) {
  const int64_t n = g.nelem();
  hipLaunchKernelGGL(k_sensitivity, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, U, nrhs, ldn, g.nx, g.ny,
                     g.nz, k, flags, xphys, dxp, phi, coef, p, emin, E0, dcdx, dvdx);
  ANNEAL_HIP(hipGetLastError());
}

void stiffness_law(const double* xphys, const double* phi, const uint8_t* flags, double E0, double emin, double p,
                   int64_t n, double* Ee, hipStream_t s) {
  hipLaunchKernelGGL(k_stiffness_law, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, xphys, phi, flags, E0,
                     emin, p, n, Ee);
  ANNEAL_HIP(hipGetLastError());
}

void design_stats(const double* xold, const double* xnew, const double* xphys, const double* phi, const double* dcx,
                  const double* dvx, const uint8_t* flags, int64_t n, double* out_dev, double* part, hipStream_t s) {
  hipLaunchKernelGGL(k_design_stats, dim3(kStatBlocks), dim3(256), 0, s, xold, xnew, xphys, phi, dcx, dvx, flags, n,
                     part);
  ANNEAL_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_design_stats_final, dim3(1), dim3(384), 0, s, part, kStatBlocks, out_dev);
  ANNEAL_HIP(hipGetLastError());
}

// Both backward passes of the sensitivities (optimize.py:142-143) in one
// launch, with the design statistics that scale the OC bracket (count, sum of
// -y1, sum of y2 over design voxels; entries 0-2 zero) when stats_dev is set.
// Radii above the ring kernels' take the two gather passes and design_stats.
void FilterContext::backward_pair(const double* d1, double* y1, const double* d2, double* y2, double* stats_dev) {
  if (R_ > kRingMaxR) {
    backward(d1, y1);
    backward(d2, y2);
    if (stats_dev) design_stats(nullptr, nullptr, nullptr, nullptr, y1, y2, flags_, g_.nelem(), stats_dev, part_, s_);
    return;
  }
  FilterArgs a = args();
  a.in = d1;
  a.in2 = d2;
  a.out = y1;
  a.out2 = y2;
  a.part = stats_dev ? part_ : nullptr;
  switch (R_) {
    case 1: hipLaunchKernelGGL((k_filter<1, kBwd2>), dim3(nblk_), dim3(FTX * FTY), 0, s_, a); break;
    case 2: hipLaunchKernelGGL((k_filter<2, kBwd2>), dim3(nblk_), dim3(FTX * FTY), 0, s_, a); break;
    case 3: hipLaunchKernelGGL((k_filter<3, kBwd2>), dim3(nblk_), dim3(FTX * FTY), 0, s_, a); break;
    case 4: hipLaunchKernelGGL((k_filter<4, kBwd2>), dim3(nblk_), dim3(FTX * FTY), 0, s_, a); break;
    case 5: hipLaunchKernelGGL((k_filter<5, kBwd2>), dim3(nblk_), dim3(FTX * FTY), 0, s_, a); break;
    default: throw std::logic_error("ring radius out of range");
  }
  ANNEAL_HIP(hipGetLastError());
  if (stats_dev) {
    hipLaunchKernelGGL(k_design_stats_final, dim3(1), dim3(384), 0, s_, part_, nblk_, stats_dev);
    ANNEAL_HIP(hipGetLastError());
  }
}

}  // namespace anneal
