// Voxel passes either side of the SIMP loop (anneal/problem.py:276-326,
// anneal/export.py:8-100, anneal/kernel.py signed distances).
//
// All of them are byte/integer or comparison work over the voxel grid and
// are HBM- or latency-bound: one thread per voxel or node, no LDS tiling.
//
// Signed distances follow the operation order of the numpy expressions they
// restate (products rounded before sums, norms summed x, y, z left to right,
// no FMA contraction), so axis-aligned boxes and spheres give the same bits as
// the reference and the same <= decisions at nodes that sit exactly on a
// threshold.  The reference evaluates the framed box and the cylinder axis
// projections with a BLAS product whose rounding depends on the host CPU;
// here they are (dx*ux + dy*uy) + dz*uz.
//
// Connected components: lock-free union-find over the face graph (a voxel
// hooks the larger of two roots under the smaller with compare-and-swap, path
// splitting on every find of the merge pass), so each component ends
// labelled with its smallest voxel index.  Parent pointers only ever decrease, which bounds every find;
// each union is also bounded and reports through an error flag.
//
// Iso level: the 40 trials of the bisection (export.py:20-27) are decided 10
// at a time from one histogram pass: the 1023 trial thresholds the next 10
// decisions can reach form a binary search tree, each voxel inside the
// current interval walks it to a leaf (voxels above it land in an extra
// slot), and the count above any trial threshold is a suffix sum over the
// slots.  4 passes instead of 40.
//
// Shortest paths: Bellman-Ford pull sweeps (each voxel recomputes its own
// distance from its neighbours when one of them changed in the last two
// sweeps, then each 64-voxel row relaxes along x in registers; no atomics on
// the distances).  The fixed point is the same
// floating-point value Dijkstra computes (left-to-right path sums,
// non-negative costs); where several predecessors give the same distance the
// smallest voxel index is taken.
#include "voxel.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <vector>

namespace anneal {
namespace {

constexpr int kThreads = 256;
inline unsigned nblk(int64_t n) { return (unsigned)std::max<int64_t>(1, (n + kThreads - 1) / kThreads); }

// Grid-shaped launches (64 x blockDim.y threads per block, one z plane per
// block layer): the kernel gets (i, j, k) without integer division.
const dim3 kBlock3(64, 4, 1);
inline dim3 grid3(int ex, int ey, int ez) { return dim3((ex + 63) / 64, (ey + 3) / 4, ez); }
struct Idx3 {
  int i, j, k;
  bool in;
};
__device__ __forceinline__ Idx3 idx3(int ex, int ey, int ez) {
  Idx3 r;
  r.i = int(blockIdx.x) * 64 + int(threadIdx.x);
  r.j = int(blockIdx.y * blockDim.y + threadIdx.y);
  r.k = int(blockIdx.z);
  r.in = r.i < ex && r.j < ey && r.k < ez;
  return r;
}

__device__ __forceinline__ double sdf_box(double qx, double qy, double qz, double hx, double hy, double hz) {
#pragma clang fp contract(off)
  const double dx = fabs(qx) - hx, dy = fabs(qy) - hy, dz = fabs(qz) - hz;
  const double mx = fmax(dx, 0.0), my = fmax(dy, 0.0), mz = fmax(dz, 0.0);
  const double outside = sqrt((mx * mx + my * my) + mz * mz);
  const double inside = fmin(fmax(fmax(dx, dy), dz), 0.0);
  return outside + inside;
}

__device__ __forceinline__ double sdf_prim(int kind, const double* __restrict__ q, double px, double py,
                                           double pz) {
#pragma clang fp contract(off)
  if (kind == kRegionBox) return sdf_box(px - q[0], py - q[1], pz - q[2], q[3], q[4], q[5]);
  if (kind == kRegionOBox) {
    const double dx = px - q[0], dy = py - q[1], dz = pz - q[2];
    const double lx = (dx * q[3] + dy * q[4]) + dz * q[5];
    const double ly = (dx * q[6] + dy * q[7]) + dz * q[8];
    double lz = (dx * q[9] + dy * q[10]) + dz * q[11];
    lz = lz - q[12];
    return sdf_box(lx, ly, lz, q[13], q[14], q[15]);
  }
  if (kind == kRegionSphere) {
    const double dx = px - q[0], dy = py - q[1], dz = pz - q[2];
    return sqrt((dx * dx + dy * dy) + dz * dz) - q[3];
  }
  // capped cylinder / pipe
  const double dx = px - q[0], dy = py - q[1], dz = pz - q[2];
  const double t = (dx * q[3] + dy * q[4]) + dz * q[5];
  const double rx = dx - t * q[3], ry = dy - t * q[4], rz = dz - t * q[5];
  const double rad = sqrt((rx * rx + ry * ry) + rz * rz);
  double dr = rad - q[7];
  if (q[8] > 0.0) dr = fmax(dr, q[8] - rad);
  const double dzc = fmax(-t, t - q[6]);
  const double m0 = fmax(dr, 0.0), m1 = fmax(dzc, 0.0);
  return sqrt(m0 * m0 + m1 * m1) + fmin(fmax(dr, dzc), 0.0);
}

__device__ __forceinline__ double anchor_sdf(const RegionTable& t, int a, double x, double y, double z) {
  double s = INFINITY;
  for (int p = t.first[a]; p < t.first[a + 1]; ++p) s = fmin(s, sdf_prim(t.kind[p], t.prm + p * kRegionParams, x, y, z));
  return s;
}

// Block-wide count of pred added to *dst with one global atomic per block
// (every thread of the block must call it; same-address global atomics are
// serialised in L2, so per-wave atomics would dominate these kernels).
__device__ __forceinline__ void block_count(bool pred, long long* dst) {
  __shared__ unsigned long long s_cnt;
  const int tid = threadIdx.x + blockDim.x * threadIdx.y;
  if (tid == 0) s_cnt = 0;
  __syncthreads();
  const unsigned long long m = __ballot(pred);
  if ((tid & 63) == 0 && m) atomicAdd(&s_cnt, (unsigned long long)__popcll(m));
  __syncthreads();
  if (tid == 0 && s_cnt) atomicAdd((unsigned long long*)dst, s_cnt);
}

template <bool NODES>
__global__ void __launch_bounds__(1024) k_region_select(int nx, int ny, int nz, double h, double ox, double oy,
                                                            double oz, RegionTable t,
                                                            const uint8_t* __restrict__ emask,
                                                            uint32_t* __restrict__ bits, long long* counts) {
#pragma clang fp contract(off)
  const int ex = NODES ? nx + 1 : nx, ey = NODES ? ny + 1 : ny, ez = NODES ? nz + 1 : nz;
  const Idx3 q = idx3(ex, ey, ez);
  uint32_t b = 0;
  if (q.in) {
    const int i = q.i, j = q.j, k = q.k;
    const int64_t p = i + int64_t(ex) * (j + int64_t(ey) * k);
    bool on = false;
    double x, y, z;
    if (NODES) {
      for (int kk = max(k - 1, 0); kk <= min(k, nz - 1) && !on; ++kk)
        for (int jj = max(j - 1, 0); jj <= min(j, ny - 1) && !on; ++jj)
          for (int ii = max(i - 1, 0); ii <= min(i, nx - 1); ++ii)
            if (emask[ii + int64_t(nx) * (jj + int64_t(ny) * kk)]) {
              on = true;
              break;
            }
      x = ox + h * double(i);
      y = oy + h * double(j);
      z = oz + h * double(k);
    } else {
      on = emask[p] != 0;
      x = ox + h * (double(i) + 0.5);
      y = oy + h * (double(j) + 0.5);
      z = oz + h * (double(k) + 0.5);
    }
    if (on)
      for (int a = 0; a < t.nanchor; ++a)
        if (anchor_sdf(t, a, x, y, z) <= t.thr[a]) b |= 1u << a;
    bits[p] = b;
  }
  __shared__ unsigned long long s_cnt[kMaxAnchors];
  const int tid = threadIdx.x + blockDim.x * threadIdx.y;
  if (tid < kMaxAnchors) s_cnt[tid] = 0;
  __syncthreads();
  for (int a = 0; a < t.nanchor; ++a) {
    const unsigned long long m = __ballot((b >> a) & 1u);
    if ((tid & 63) == 0 && m) atomicAdd(&s_cnt[a], (unsigned long long)__popcll(m));
  }
  __syncthreads();
  if (tid < t.nanchor && s_cnt[tid]) atomicAdd((unsigned long long*)(counts + tid), s_cnt[tid]);
}

__global__ void __launch_bounds__(kThreads) k_region_sdf(int64_t n, const double* __restrict__ pts, RegionTable t,
                                                         double* __restrict__ out) {
  const int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const double x = pts[3 * p], y = pts[3 * p + 1], z = pts[3 * p + 2];
  for (int a = 0; a < t.nanchor; ++a) out[a * n + p] = anchor_sdf(t, a, x, y, z);
}

__global__ void __launch_bounds__(kThreads) k_boundary_fields(int64_t nnode, const uint32_t* __restrict__ bits,
                                                              int nanchor, const uint8_t* __restrict__ dofmask,
                                                              const double* __restrict__ load, int ncase,
                                                              uint8_t* __restrict__ fixed_bits,
                                                              double* __restrict__ F) {
#pragma clang fp contract(off)
  const int64_t n = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (n >= nnode) return;
  const uint32_t b = bits[n];
  uint8_t fb = 0;
  for (int a = 0; a < nanchor; ++a)
    if ((b >> a) & 1u) fb |= dofmask[a];
  fixed_bits[n] = fb;
  for (int c = 0; c < ncase; ++c)
    for (int d = 0; d < 3; ++d) {
      double acc = 0.0;
      for (int a = 0; a < nanchor; ++a)
        if ((b >> a) & 1u) acc += load[(int64_t(c) * nanchor + a) * 3 + d];
      F[int64_t(c) * 3 * nnode + 3 * n + d] = acc;
    }
}

__global__ void __launch_bounds__(kThreads) k_gather_node_bits(int nx, int ny, int nz,
                                                               const uint32_t* __restrict__ nb,
                                                               const uint8_t* __restrict__ emask,
                                                               uint32_t* __restrict__ eb) {
  const Idx3 q = idx3(nx, ny, nz);
  if (!q.in) return;
  const int64_t i = q.i, j = q.j, k = q.k;
  const int64_t e = i + int64_t(nx) * (j + int64_t(ny) * k);
  uint32_t b = 0;
  if (emask[e]) {
    const int64_t sx = nx + 1, sxy = int64_t(nx + 1) * (ny + 1);
    const int64_t n0 = i + sx * j + sxy * k;
    b = nb[n0] | nb[n0 + 1] | nb[n0 + sx] | nb[n0 + sx + 1] | nb[n0 + sxy] | nb[n0 + sxy + 1] | nb[n0 + sxy + sx] |
        nb[n0 + sxy + sx + 1];
  }
  eb[e] = b;
}

// ---- connected components ---------------------------------------------------
__device__ __forceinline__ int ld_par(const int32_t* p, int x) {
  return __hip_atomic_load(p + x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_par(int32_t* p, int x, int v) {
  __hip_atomic_store(p + x, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Root of x.  parent[y] < y for every non-root y, so the walk strictly
// decreases and ends; path splitting points each visited node at its
// grandparent (still an ancestor, still smaller).
__device__ int find_root(int32_t* par, int x) {
  int cur = x, nxt = ld_par(par, x);
  while (cur != nxt) {
    const int gp = ld_par(par, nxt);
    if (gp != nxt) st_par(par, cur, gp);
    cur = nxt;
    nxt = gp;
  }
  return cur;
}

// Root of x without writes (the labelling pass writes only final roots).
__device__ int find_root_ro(const int32_t* par, int x) {
  int p = ld_par(par, x);
  while (p != x) {
    x = p;
    p = ld_par(par, x);
  }
  return x;
}

constexpr int kUnionBound = 1 << 24;

__device__ void unite(int32_t* par, int a, int b, int* err) {
  for (int it = 0; it < kUnionBound; ++it) {
    a = find_root(par, a);
    b = find_root(par, b);
    if (a == b) return;
    if (a > b) {
      const int t = a;
      a = b;
      b = t;
    }
    if (atomicCAS(par + b, b, a) == b) return;  // b was still a root: now under a
  }
  atomicOr(err, 1);
}

__global__ void __launch_bounds__(kThreads) k_cc_init(int64_t ne, const uint8_t* __restrict__ mask,
                                                      int32_t* __restrict__ par) {
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e < ne) par[e] = mask[e] ? int32_t(e) : -1;
}

__global__ void __launch_bounds__(kThreads) k_cc_merge(int nx, int ny, int nz, const uint8_t* __restrict__ mask,
                                                       int32_t* par, int* err) {
  const Idx3 q = idx3(nx, ny, nz);
  if (!q.in) return;
  const int64_t i = q.i, j = q.j, k = q.k, sxy = int64_t(nx) * ny;
  const int64_t e = i + nx * j + sxy * k;
  if (!mask[e]) return;
  if (i + 1 < nx && mask[e + 1]) unite(par, int(e), int(e + 1), err);
  if (j + 1 < ny && mask[e + nx]) unite(par, int(e), int(e + nx), err);
  if (k + 1 < nz && mask[e + sxy]) unite(par, int(e), int(e + sxy), err);
}

// Labelling pass (after the merge pass, so every root is final): each voxel
// finds its root without writes, then points every node of its path at that
// root.  Only final roots are ever written, so concurrent walks stay valid and
// no label is replaced by a non-root.
__global__ void __launch_bounds__(1024) k_cc_compress(int64_t ne, int32_t* par, long long* nroot) {
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  bool root = false;
  if (e < ne && ld_par(par, int(e)) >= 0) {
    const int r = find_root_ro(par, int(e));
    int cur = int(e);
    while (cur != r) {
      const int nxt = ld_par(par, cur);
      if (nxt != r) st_par(par, cur, r);
      cur = nxt;
    }
    st_par(par, int(e), r);
    root = r == int(e);
  }
  block_count(root, nroot);
}

__global__ void k_zero_ll(long long* p, int n) {
  if (int(threadIdx.x) < n) p[threadIdx.x] = 0;
}

__global__ void k_zero_i(int* p) {
  if (threadIdx.x == 0) *p = 0;
}

__global__ void __launch_bounds__(kThreads) k_fill_u8(int64_t n, uint8_t* p, uint8_t v) {
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e < n) p[e] = v;
}

__global__ void __launch_bounds__(kThreads) k_link_mark(int64_t ne, const uint8_t* __restrict__ mask,
                                                        const int32_t* __restrict__ lab,
                                                        const uint32_t* __restrict__ eb, uint32_t seed,
                                                        uint8_t* keep) {
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e < ne && mask[e] && (eb[e] & seed)) keep[lab[e]] = 1;
}

// work[0]: marked components, work[1]: OR of ebits over their voxels.
__global__ void __launch_bounds__(1024) k_link_apply(int64_t ne, const uint8_t* __restrict__ mask,
                                                         const int32_t* __restrict__ lab,
                                                         const uint32_t* __restrict__ eb,
                                                         const uint8_t* __restrict__ keep, uint8_t* out_mask,
                                                         long long* work) {
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  bool m = false;
  uint32_t b = 0;
  if (e < ne) {
    m = mask[e] && keep[lab[e]];
    if (out_mask) out_mask[e] = m ? 1 : 0;
    if (m) b = eb[e];
  }
  block_count(m && lab[e] == int32_t(e), work);
  for (int off = 32; off > 0; off >>= 1) b |= __shfl_xor(b, off);
  __shared__ unsigned s_or;
  if (threadIdx.x == 0) s_or = 0;
  __syncthreads();
  if ((threadIdx.x & 63) == 0 && b) atomicOr(&s_or, b);
  __syncthreads();
  if (threadIdx.x == 0 && s_or) atomicOr((unsigned long long*)(work + 1), (unsigned long long)s_or);
}

__global__ void __launch_bounds__(kThreads) k_threshold(int64_t n, const double* __restrict__ f, double t,
                                                        uint8_t* __restrict__ mask) {
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e < n) mask[e] = f[e] > t ? 1 : 0;
}

__global__ void __launch_bounds__(1024) k_count_above(int64_t n, const double* __restrict__ f, double t,
                                                          long long* cnt) {
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  block_count(e < n && f[e] > t, cnt);
}

__global__ void __launch_bounds__(kThreads) k_output_field(int64_t n, const double* __restrict__ x,
                                                           const double* __restrict__ fill,
                                                           const uint8_t* __restrict__ active,
                                                           const double* __restrict__ solid,
                                                           const double* __restrict__ vd,
                                                           const uint8_t* __restrict__ frozen,
                                                           double* __restrict__ f) {
#pragma clang fp contract(off)
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= n) return;
  double v = x[e] * fill[e];
  v = fmax(v, active[e] ? solid[e] : 0.0);
  v = fmin(v, 1.0 - vd[e]);
  if (frozen[e]) v = fmax(v, fill[e]);
  f[e] = v;
}

// ---- iso level ------------------------------------------------------------
constexpr int kLevelDepth = 10;                   // decisions per pass
constexpr int kLevelLeaves = 1 << kLevelDepth;    // 1024 slots
constexpr int kLevelPasses = 4;                   // 40 decisions
constexpr int kSumBlocks = 1024;

struct LevelState {
  double lo, hi, target, pad;
};

// hist[r][kLevelLeaves] is the "over" slot: values above the current
// interval, which exceed every trial threshold of the pass.  kHistCopies
// copies of the histogram (block b adds into copy b % kHistCopies) spread the
// global atomics over more cache lines.
constexpr int kHistCopies = 64;
constexpr int kHistStride = kLevelLeaves + 64;
struct LevelWork {
  LevelState st;
  double part[kSumBlocks];
  unsigned hist[kHistCopies][kHistStride];
};

// Deterministic FP64 sum of x*fill: fixed block partials, then one block.
__global__ void __launch_bounds__(kThreads) k_level_sum(int64_t n, const double* __restrict__ x,
                                                        const double* __restrict__ fill, LevelWork* w) {
  __shared__ double s[kThreads];
  double acc = 0.0;
  for (int64_t e = int64_t(blockIdx.x) * kThreads + threadIdx.x; e < n; e += int64_t(gridDim.x) * kThreads)
    acc += x[e] * fill[e];
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int o = kThreads / 2; o > 0; o >>= 1) {
    if (int(threadIdx.x) < o) s[threadIdx.x] += s[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) w->part[blockIdx.x] = s[0];
}

__global__ void __launch_bounds__(kThreads) k_level_init(LevelWork* w) {
  __shared__ double s[kThreads];
  double acc = 0.0;
  for (int b = threadIdx.x; b < kSumBlocks; b += kThreads) acc += w->part[b];
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int o = kThreads / 2; o > 0; o >>= 1) {
    if (int(threadIdx.x) < o) s[threadIdx.x] += s[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    w->st.target = s[0];
    w->st.lo = 0.05;
    w->st.hi = 0.95;
  }
  for (int b = threadIdx.x; b < kHistCopies * kHistStride; b += kThreads) (&w->hist[0][0])[b] = 0;
}

// Trial threshold of heap node k (root 1) at depth d below the interval (lo, hi):
// the bisection's own midpoints along the path to k.
__device__ __forceinline__ double tree_threshold(double lo, double hi, int k, int d) {
#pragma clang fp contract(off)
  for (int l = d - 1; l >= 0; --l) {
    const double t = 0.5 * (lo + hi);
    if ((k >> l) & 1) lo = t;
    else hi = t;
  }
  return 0.5 * (lo + hi);
}

// Values <= lo exceed no trial threshold of the pass and are skipped; values
// > hi exceed all of them (counted per thread, then per block); the rest walk
// the tree to their slot.
__global__ void __launch_bounds__(kThreads) k_level_hist(int64_t n, const double* __restrict__ f, LevelWork* w) {
  __shared__ double thr[kLevelLeaves];
  __shared__ unsigned hist[kLevelLeaves + 1];
  const double lo = w->st.lo, hi = w->st.hi;
  for (int k = 1 + threadIdx.x; k < kLevelLeaves; k += kThreads) thr[k] = tree_threshold(lo, hi, k, 31 - __clz(k));
  for (int b = threadIdx.x; b <= kLevelLeaves; b += kThreads) hist[b] = 0;
  __syncthreads();
  unsigned over = 0;
  for (int64_t e = int64_t(blockIdx.x) * kThreads + threadIdx.x; e < n; e += int64_t(gridDim.x) * kThreads) {
    const double v = f[e];
    if (!(v > lo)) continue;
    if (v > hi) {
      ++over;
      continue;
    }
    int k = 1;
#pragma unroll
    for (int l = 0; l < kLevelDepth; ++l) k = 2 * k + (v > thr[k] ? 1 : 0);
    atomicAdd(&hist[k - kLevelLeaves], 1u);
  }
  for (int off = 32; off > 0; off >>= 1) over += __shfl_xor(over, off);
  if ((threadIdx.x & 63) == 0 && over) atomicAdd(&hist[kLevelLeaves], over);
  __syncthreads();
  unsigned* dst = w->hist[blockIdx.x % kHistCopies];
  for (int b = threadIdx.x; b <= kLevelLeaves; b += kThreads)
    if (hist[b]) atomicAdd(&dst[b], hist[b]);
}

// Replays kLevelDepth decisions from the histogram (count above the trial of
// in-order rank r = slots r+1 .. over), narrows (lo, hi] to the final slot
// and clears the histogram for the next pass.
__global__ void __launch_bounds__(kLevelLeaves) k_level_walk(LevelWork* w) {
#pragma clang fp contract(off)
  // suffix sums suf[s] = count in slots >= s (s <= kLevelLeaves), by doubling steps in LDS
  __shared__ unsigned long long buf[2][kLevelLeaves + 1];
  const int t = threadIdx.x;
  unsigned long long h = 0, ho = 0;
  for (int r = 0; r < kHistCopies; ++r) {
    h += w->hist[r][t];
    if (t == 0) ho += w->hist[r][kLevelLeaves];
  }
  buf[0][t] = h;
  if (t == 0) buf[0][kLevelLeaves] = ho;
  __syncthreads();
  int cur = 0;
  for (int off = 1; off <= kLevelLeaves; off <<= 1) {
    buf[cur ^ 1][t] = buf[cur][t] + (t + off <= kLevelLeaves ? buf[cur][t + off] : 0ull);
    if (t == 0) buf[cur ^ 1][kLevelLeaves] = buf[cur][kLevelLeaves];
    __syncthreads();
    cur ^= 1;
  }
  const unsigned long long* suf = buf[cur];
  if (t == 0) {
    double lo = w->st.lo, hi = w->st.hi;
    const double target = w->st.target;
    int k = 1;
    for (int d = 0; d < kLevelDepth; ++d) {
      const double th = 0.5 * (lo + hi);
      const int rank = (2 * (k - (1 << d)) + 1) * (1 << (kLevelDepth - 1 - d)) - 1;
      const double v = double(suf[rank + 1]);
      if (v > target) {
        lo = th;
        k = 2 * k + 1;
      } else {
        hi = th;
        k = 2 * k;
      }
    }
    w->st.lo = lo;
    w->st.hi = hi;
  }
  for (int r = 0; r < kHistCopies; ++r) {
    w->hist[r][t] = 0;
    if (t == 0) w->hist[r][kLevelLeaves] = 0;
  }
}

// ---- shortest paths -------------------------------------------------------
__device__ __forceinline__ double path_cost(const double* f, int64_t u, double need) {
#pragma clang fp contract(off)
  return fmax(0.0, need - f[u]) + 1e-3;
}

__device__ __forceinline__ double ld_dist(const double* d, int64_t i) {
  return __longlong_as_double(
      (long long)__hip_atomic_load((const unsigned long long*)(d + i), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
}

__global__ void __launch_bounds__(kThreads) k_sssp_init(int64_t ne, const uint8_t* __restrict__ act,
                                                        const uint32_t* __restrict__ eb, uint32_t src,
                                                        double* __restrict__ dist, int32_t* __restrict__ pred,
                                                        int32_t* __restrict__ stamp) {
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  const bool s = act[e] && (eb[e] & src);
  dist[e] = s ? 0.0 : INFINITY;
  pred[e] = -1;
  stamp[e] = s ? 0 : -2;
}

// One pull sweep: an active voxel whose neighbours changed in this sweep or
// the previous one takes min(dist, dist[v] + cost(u)) over its active
// neighbours v; then each wave (64 consecutive voxels along x) relaxes its
// row in register, forward and backward, one edge at a time, so a distance
// crosses the whole row within the sweep.  Every update is one exact
// relaxation d[u] = min(d[u], d[v] + cost(u)), so the fixed point is the
// same.  Only the owner writes dist[u] (values only decrease; a neighbour
// read mid-update is an older, larger distance and the voxel is revisited
// next sweep through its stamp).
__global__ void __launch_bounds__(kThreads) k_sssp_sweep(int nx, int ny, int nz, const uint8_t* __restrict__ act,
                                                         const double* __restrict__ f, double need, double* dist,
                                                         int32_t* stamp, int* flag, int sweep) {
#pragma clang fp contract(off)
  const Idx3 q3 = idx3(nx, ny, nz);
  const int i = q3.i, j = q3.j, k = q3.k;
  const int lane = threadIdx.x;  // blockDim.x == 64: one wave per row segment
  const int64_t sxy = int64_t(nx) * ny;
  const int64_t u = q3.in ? i + int64_t(nx) * j + sxy * k : 0;
  const bool a = q3.in && act[u];
  int64_t nb[6];
  int cnt = 0;
  bool fresh = false;
  if (a) {
    if (i > 0) nb[cnt++] = u - 1;
    if (i + 1 < nx) nb[cnt++] = u + 1;
    if (j > 0) nb[cnt++] = u - nx;
    if (j + 1 < ny) nb[cnt++] = u + nx;
    if (k > 0) nb[cnt++] = u - sxy;
    if (k + 1 < nz) nb[cnt++] = u + sxy;
    for (int q = 0; q < cnt; ++q)
      fresh = fresh || __hip_atomic_load(stamp + nb[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= sweep - 1;
  }
  if (__ballot(fresh) == 0) return;  // wave-uniform
  const double cu = a ? path_cost(f, u, need) : 0.0;
  const double du = a ? ld_dist(dist, u) : INFINITY;
  double best = du;
  if (fresh)
    for (int q = 0; q < cnt; ++q) {
      const int64_t v = nb[q];
      if (!act[v]) continue;
      const double c = ld_dist(dist, v) + cu;
      if (c < best) best = c;
    }
  // in-row relaxation; inactive lanes hold +inf and never change
  for (int it = 0; it < 63; ++it) {
    const double from = __shfl_up(best, 1);
    const double c = from + cu;
    const bool imp = a && lane > 0 && c < best;
    if (imp) best = c;
    if (__ballot(imp) == 0) break;
  }
  for (int it = 0; it < 63; ++it) {
    const double from = __shfl_down(best, 1);
    const double c = from + cu;
    const bool imp = a && lane < 63 && c < best;
    if (imp) best = c;
    if (__ballot(imp) == 0) break;
  }
  if (a && best < du) {
    __hip_atomic_store((unsigned long long*)(dist + u), (unsigned long long)__double_as_longlong(best),
                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(stamp + u, sweep, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    // one atomic per sweep and L2 slice in practice: later changers see the flag already set
    if (__hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < sweep + 1) atomicMax(flag, sweep + 1);
  }
}

__global__ void __launch_bounds__(kThreads) k_sssp_pred(int nx, int ny, int nz, const uint8_t* __restrict__ act,
                                                        const double* __restrict__ f, double need,
                                                        const uint32_t* __restrict__ eb, uint32_t src,
                                                        const double* __restrict__ dist,
                                                        int32_t* __restrict__ pred) {
#pragma clang fp contract(off)
  const Idx3 q3 = idx3(nx, ny, nz);
  if (!q3.in) return;
  const int64_t i = q3.i, j = q3.j, k = q3.k, sxy = int64_t(nx) * ny;
  const int64_t u = i + nx * j + sxy * k;
  int32_t p = -1;
  const double du = dist[u];
  if (act[u] && !(eb[u] & src) && du < INFINITY) {
    const double cu = path_cost(f, u, need);
    // neighbours in increasing index order: the first match is the smallest
    const int64_t cand[6] = {k > 0 ? u - sxy : -1,     j > 0 ? u - nx : -1,      i > 0 ? u - 1 : -1,
                             i + 1 < nx ? u + 1 : -1, j + 1 < ny ? u + nx : -1, k + 1 < nz ? u + sxy : -1};
    for (int q = 0; q < 6; ++q) {
      const int64_t v = cand[q];
      if (v < 0 || !act[v]) continue;
      if (dist[v] + cu == du) {
        p = int32_t(v);
        break;
      }
    }
  }
  pred[u] = p;
}

// work[0]: min distance bits over the group, work[1]: smallest index reaching it.
__global__ void __launch_bounds__(kThreads) k_group_min(int64_t ne, const uint8_t* __restrict__ act,
                                                        const uint32_t* __restrict__ eb, uint32_t gbit,
                                                        const double* __restrict__ dist, long long* work) {
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  unsigned long long d = ~0ull;
  if (e < ne && act[e] && (eb[e] & gbit)) d = (unsigned long long)__double_as_longlong(dist[e]);
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned long long o = __shfl_xor(d, off);
    d = o < d ? o : d;
  }
  if ((threadIdx.x & 63) == 0 && d != ~0ull) atomicMin((unsigned long long*)work, d);
}

__global__ void __launch_bounds__(kThreads) k_group_argmin(int64_t ne, const uint8_t* __restrict__ act,
                                                           const uint32_t* __restrict__ eb, uint32_t gbit,
                                                           const double* __restrict__ dist, long long* work) {
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e < ne && act[e] && (eb[e] & gbit) &&
      (unsigned long long)__double_as_longlong(dist[e]) == (unsigned long long)work[0])
    atomicMin(work + 1, (long long)e);
}

// One thread: follow pred from the start voxel; at most ne steps.
__global__ void k_trace(int64_t ne, const int32_t* __restrict__ pred, const double* __restrict__ dist,
                        int32_t* __restrict__ path, long long* work) {
  const long long start = work[1];
  long long len = 0;
  if (start >= 0 && start < ne && dist[start] < INFINITY) {
    long long v = start;
    while (v >= 0 && len < ne) {
      path[len++] = int32_t(v);
      v = pred[v];
    }
  }
  work[2] = len;
}

__global__ void __launch_bounds__(kThreads) k_raise_ring(int nx, int ny, int64_t ne, const int32_t* __restrict__ path,
                                                         int64_t len, const uint8_t* __restrict__ act,
                                                         const double* __restrict__ fill, double val_need,
                                                         double* f) {
#pragma clang fp contract(off)
  const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= 6 * len) return;
  const int64_t sxy = int64_t(nx) * ny;
  const int64_t off[6] = {1, -1, nx, -int64_t(nx), sxy, -sxy};
  const int64_t r = int64_t(path[t / 6]) + off[t % 6];
  if (r < 0 || r >= ne || !act[r]) return;
  f[r] = fmax(f[r], fmin(val_need, fill[r]));
}

__global__ void __launch_bounds__(kThreads) k_raise_path(const int32_t* __restrict__ path, int64_t len,
                                                         const double* __restrict__ fill, double* f) {
  const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= len) return;
  const int64_t v = path[t];
  f[v] = fmax(f[v], fmin(1.0, fill[v]));
}

}  // namespace

// ---------------------------------------------------------------------------
void region_select(const GridDims& g, double h, const double* origin, const RegionTable& t, bool at_nodes,
                   const uint8_t* emask, uint32_t* bits, long long* counts, hipStream_t s) {
  ANNEAL_CHECK(t.nanchor >= 0 && t.nanchor <= kMaxAnchors, "at most 32 anchors");
  k_zero_ll<<<1, 64, 0, s>>>(counts, kMaxAnchors);
  // 64 x 16 threads per block: one global atomic per block and anchor
  const dim3 blk(64, 16, 1);
  if (at_nodes)
    k_region_select<true><<<dim3((g.nx + 64) / 64, (g.ny + 16) / 16, g.nz + 1), blk, 0, s>>>(g.nx, g.ny, g.nz, h, origin[0], origin[1], origin[2], t, emask,
                                                       bits, counts);
  else
    k_region_select<false><<<dim3((g.nx + 63) / 64, (g.ny + 15) / 16, g.nz), blk, 0, s>>>(g.nx, g.ny, g.nz, h, origin[0], origin[1], origin[2], t,
                                                        emask, bits, counts);
  ANNEAL_HIP(hipGetLastError());
}

void region_sdf(int64_t n, const double* pts, const RegionTable& t, double* out, hipStream_t s) {
  if (n == 0) return;
  k_region_sdf<<<nblk(n), kThreads, 0, s>>>(n, pts, t, out);
  ANNEAL_HIP(hipGetLastError());
}

void boundary_fields(const GridDims& g, const uint32_t* node_bits, int nanchor, const uint8_t* dofmask,
                     const double* load, int ncase, uint8_t* fixed_bits, double* F, hipStream_t s) {
  k_boundary_fields<<<nblk(g.nnode()), kThreads, 0, s>>>(g.nnode(), node_bits, nanchor, dofmask, load, ncase,
                                                         fixed_bits, F);
  ANNEAL_HIP(hipGetLastError());
}

void gather_node_bits(const GridDims& g, const uint32_t* node_bits, const uint8_t* emask, uint32_t* ebits,
                      hipStream_t s) {
  k_gather_node_bits<<<grid3(g.nx, g.ny, g.nz), kBlock3, 0, s>>>(g.nx, g.ny, g.nz, node_bits, emask, ebits);
  ANNEAL_HIP(hipGetLastError());
}

int64_t components(const GridDims& g, const uint8_t* mask, int32_t* label, long long* counter, hipStream_t s) {
  ANNEAL_CHECK(g.nelem() < (int64_t(1) << 31), "grid too large for 32-bit labels");
  const int64_t ne = g.nelem();
  k_zero_ll<<<1, 64, 0, s>>>(counter, 2);
  k_cc_init<<<nblk(ne), kThreads, 0, s>>>(ne, mask, label);
  k_cc_merge<<<grid3(g.nx, g.ny, g.nz), kBlock3, 0, s>>>(g.nx, g.ny, g.nz, mask, label, (int*)(counter + 1));
  k_cc_compress<<<unsigned((ne + 1023) / 1024), 1024, 0, s>>>(ne, label, counter);
  ANNEAL_HIP(hipGetLastError());
  long long host[2] = {0, 0};
  ANNEAL_HIP(hipMemcpyAsync(host, counter, 2 * sizeof(long long), hipMemcpyDeviceToHost, s));
  ANNEAL_HIP(hipStreamSynchronize(s));
  if (host[1]) throw std::runtime_error("connected components: a union step exceeded its bound");
  return host[0];
}

int64_t link_components(const GridDims& g, const uint8_t* mask, const int32_t* label, const uint32_t* ebits,
                        uint32_t seed, uint8_t* keep, uint8_t* out_mask, uint32_t* linked, long long* counter,
                        hipStream_t s) {
  const int64_t ne = g.nelem();
  k_zero_ll<<<1, 64, 0, s>>>(counter, 2);
  k_fill_u8<<<nblk(ne), kThreads, 0, s>>>(ne, keep, 0);
  k_link_mark<<<nblk(ne), kThreads, 0, s>>>(ne, mask, label, ebits, seed, keep);
  k_link_apply<<<unsigned((ne + 1023) / 1024), 1024, 0, s>>>(ne, mask, label, ebits, keep, out_mask, counter);
  ANNEAL_HIP(hipGetLastError());
  long long host[2] = {0, 0};
  ANNEAL_HIP(hipMemcpyAsync(host, counter, 2 * sizeof(long long), hipMemcpyDeviceToHost, s));
  ANNEAL_HIP(hipStreamSynchronize(s));
  if (linked) *linked = uint32_t(host[1]);
  return host[0];
}

void threshold_mask(int64_t n, const double* f, double t, uint8_t* mask, hipStream_t s) {
  if (n == 0) return;
  k_threshold<<<nblk(n), kThreads, 0, s>>>(n, f, t, mask);
  ANNEAL_HIP(hipGetLastError());
}

int64_t count_above(int64_t n, const double* f, double t, long long* counter, hipStream_t s) {
  k_zero_ll<<<1, 64, 0, s>>>(counter, 1);
  if (n > 0) k_count_above<<<unsigned((n + 1023) / 1024), 1024, 0, s>>>(n, f, t, counter);
  ANNEAL_HIP(hipGetLastError());
  long long host = 0;
  ANNEAL_HIP(hipMemcpyAsync(&host, counter, sizeof(long long), hipMemcpyDeviceToHost, s));
  ANNEAL_HIP(hipStreamSynchronize(s));
  return host;
}

void output_field(int64_t n, const double* x, const double* fill, const uint8_t* active, const double* solid_fill,
                  const double* void_fill, const uint8_t* frozen, double* f, hipStream_t s) {
  if (n == 0) return;
  k_output_field<<<nblk(n), kThreads, 0, s>>>(n, x, fill, active, solid_fill, void_fill, frozen, f);
  ANNEAL_HIP(hipGetLastError());
}

size_t choose_level_work_bytes() { return sizeof(LevelWork); }

double choose_level(int64_t n, const double* f, const double* x, const double* fill, void* work, double* target_out,
                    hipStream_t s) {
  LevelWork* w = (LevelWork*)work;
  // every one of the kSumBlocks partials is written (blocks past the data write 0)
  k_level_sum<<<kSumBlocks, kThreads, 0, s>>>(n, x, fill, w);
  const unsigned hb = (unsigned)std::min<int64_t>(1024, std::max<int64_t>(1, (n + kThreads - 1) / kThreads));
  k_level_init<<<1, kThreads, 0, s>>>(w);
  for (int p = 0; p < kLevelPasses; ++p) {
    k_level_hist<<<hb, kThreads, 0, s>>>(n, f, w);
    k_level_walk<<<1, kLevelLeaves, 0, s>>>(w);
  }
  ANNEAL_HIP(hipGetLastError());
  LevelState st;
  ANNEAL_HIP(hipMemcpyAsync(&st, &w->st, sizeof(st), hipMemcpyDeviceToHost, s));
  ANNEAL_HIP(hipStreamSynchronize(s));
  if (target_out) *target_out = st.target;
  const double mid = 0.5 * (st.lo + st.hi);
  return std::min(std::max(mid, 0.3), 0.7);
}

int sssp(const GridDims& g, const uint8_t* active, const double* f, double need, const uint32_t* ebits,
         uint32_t source_groups, double* dist, int32_t* pred, int32_t* stamp, int* flag, int max_sweeps,
         hipStream_t s) {
  ANNEAL_CHECK(g.nelem() < (int64_t(1) << 31), "grid too large for 32-bit voxel indices");
  const int64_t ne = g.nelem();
  k_sssp_init<<<nblk(ne), kThreads, 0, s>>>(ne, active, ebits, source_groups, dist, pred, stamp);
  k_zero_i<<<1, 64, 0, s>>>(flag);
  ANNEAL_HIP(hipGetLastError());
  constexpr int kBatch = 32;
  int sweep = 0;
  bool done = false;
  while (!done && sweep < max_sweeps) {
    const int upto = std::min(sweep + kBatch, max_sweeps);
    for (; sweep < upto; ++sweep)
      k_sssp_sweep<<<grid3(g.nx, g.ny, g.nz), kBlock3, 0, s>>>(g.nx, g.ny, g.nz, active, f, need, dist, stamp, flag, sweep);
    ANNEAL_HIP(hipGetLastError());
    int last = 0;
    ANNEAL_HIP(hipMemcpyAsync(&last, flag, sizeof(int), hipMemcpyDeviceToHost, s));
    ANNEAL_HIP(hipStreamSynchronize(s));
    done = last < sweep;  // the last sweep improved nothing
  }
  if (!done) throw std::runtime_error("shortest paths did not settle within the sweep bound");
  k_sssp_pred<<<grid3(g.nx, g.ny, g.nz), kBlock3, 0, s>>>(g.nx, g.ny, g.nz, active, f, need, ebits, source_groups, dist, pred);
  ANNEAL_HIP(hipGetLastError());
  return sweep;
}

int64_t trace_path(const GridDims& g, const uint8_t* active, const uint32_t* ebits, int group, const double* dist,
                   const int32_t* pred, int32_t* path, long long* work, int64_t* start, double* dstart,
                   hipStream_t s) {
  ANNEAL_CHECK(group >= 0 && group < kMaxAnchors, "group index out of range");
  const int64_t ne = g.nelem();
  const uint32_t gbit = 1u << group;
  const long long init[3] = {(long long)0x7FFFFFFFFFFFFFFFLL, (long long)0x7FFFFFFFFFFFFFFFLL, 0};
  ANNEAL_HIP(hipMemcpyAsync(work, init, sizeof(init), hipMemcpyHostToDevice, s));
  k_group_min<<<nblk(ne), kThreads, 0, s>>>(ne, active, ebits, gbit, dist, work);
  k_group_argmin<<<nblk(ne), kThreads, 0, s>>>(ne, active, ebits, gbit, dist, work);
  k_trace<<<1, 1, 0, s>>>(ne, pred, dist, path, work);
  ANNEAL_HIP(hipGetLastError());
  long long host[3];
  ANNEAL_HIP(hipMemcpyAsync(host, work, sizeof(host), hipMemcpyDeviceToHost, s));
  ANNEAL_HIP(hipStreamSynchronize(s));
  const bool empty = host[1] == (long long)0x7FFFFFFFFFFFFFFFLL;
  *start = empty ? -1 : host[1];
  if (dstart) {
    double d;
    std::memcpy(&d, &host[0], sizeof(d));
    *dstart = empty ? std::numeric_limits<double>::infinity() : d;
  }
  return host[2];
}

void raise_path(const GridDims& g, const int32_t* path, int64_t len, const uint8_t* active, const double* fill,
                double need, double* f, hipStream_t s) {
  if (len <= 0) return;
  k_raise_ring<<<nblk(6 * len), kThreads, 0, s>>>(g.nx, g.ny, g.nelem(), path, len, active, fill, need + 0.05, f);
  k_raise_path<<<nblk(len), kThreads, 0, s>>>(path, len, fill, f);
  ANNEAL_HIP(hipGetLastError());
}

}  // namespace anneal
