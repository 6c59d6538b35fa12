// Vector kernels with deterministic FP64 reductions (fixed grid, fixed tree).
#include "common.hpp"
#include "blas.hpp"

#include <algorithm>

namespace anneal {

constexpr int kRedThreads = 256;

template <typename TA, typename TB>
__global__ __launch_bounds__(kRedThreads) void k_dot_partial(const TA* __restrict__ a, const TB* __restrict__ b,
                                                             int64_t n, int64_t ld, double* __restrict__ part,
                                                             const double* __restrict__ skip) {
  const int r = blockIdx.y;
  if (rhs_skipped(skip, r)) return;
  const TA* pa = a + (int64_t)3 * r * ld;
  const TB* pb = b + (int64_t)3 * r * ld;
  double s = 0.0;
  // four grid-stride entries per pass with their loads issued first (clamped
  // addresses); the products are summed in the entry order of a plain loop
  constexpr int U = 4;
  const int64_t stride = (int64_t)gridDim.x * kRedThreads;
  for (int c = 0; c < 3; ++c) {
    const TA* qa = pa + c * ld;
    const TB* qb = pb + c * ld;
    for (int64_t i0 = blockIdx.x * (int64_t)kRedThreads + threadIdx.x; i0 < n; i0 += U * stride) {
      double va[U], vb[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int64_t i = i0 + u * stride < n ? i0 + u * stride : n - 1;
        va[u] = (double)qa[i];
        vb[u] = (double)qb[i];
      }
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (i0 + u * stride < n) s += va[u] * vb[u];
    }
  }
  __shared__ double sh[kRedThreads / 64];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < kRedThreads / 64; ++w) t += sh[w];
    part[(int64_t)r * gridDim.x + blockIdx.x] = t;
  }
}

// Sum nblk partials per rhs in a fixed order; out[r] = sum (or out[r] = scale*sum).
__global__ __launch_bounds__(kRedThreads) void k_finalize(const double* __restrict__ part, int nblk,
                                                          double* __restrict__ out, double* __restrict__ out_host) {
  const int r = blockIdx.x;
  double s = 0.0;
  for (int i = threadIdx.x; i < nblk; i += kRedThreads) s += part[(int64_t)r * nblk + i];
  __shared__ double sh[kRedThreads / 64];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < kRedThreads / 64; ++w) t += sh[w];
    out[r] = t;
    if (out_host) out_host[r] = t;  // pinned host memory, read after the caller's wait
  }
}

template <typename TA, typename TB>
void dot(const TA* a, const TB* b, int64_t n, int64_t ld, int nrhs, double* part, double* out, hipStream_t s,
         double* out_host) {
  hipLaunchKernelGGL((k_dot_partial<TA, TB>), dim3(kDotBlocks, nrhs), dim3(kRedThreads), 0, s, a, b, n, ld, part,
                     (const double*)nullptr);
  ANNEAL_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_finalize, dim3(nrhs), dim3(kRedThreads), 0, s, part, kDotBlocks, out, out_host);
  ANNEAL_HIP(hipGetLastError());
}
template void dot<float, float>(const float*, const float*, int64_t, int64_t, int, double*, double*, hipStream_t,
                                double*);
template void dot<double, double>(const double*, const double*, int64_t, int64_t, int, double*, double*, hipStream_t,
                                  double*);

void finalize(const double* part, int nblk, int nrhs, double* out, hipStream_t s, double* out_host) {
  hipLaunchKernelGGL(k_finalize, dim3(nrhs), dim3(kRedThreads), 0, s, part, nblk, out, out_host);
  ANNEAL_HIP(hipGetLastError());
}

// x = M * omega * dinv * b  (first Jacobi sweep from a zero guess)
__global__ void k_jacobi0(const float* __restrict__ b, float* __restrict__ x, const float* __restrict__ dinv,
                          const uint8_t* __restrict__ mask, float omega, int64_t n, int64_t ld,
                          const double* __restrict__ skip) {
  const int r = blockIdx.y;
  if (rhs_skipped(skip, r)) return;
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float d = omega * dinv[i];
  const uint8_t mk = mask[i];
  for (int c = 0; c < 3; ++c) {
    const int64_t o = (int64_t)(3 * r + c) * ld + i;
    x[o] = ((mk >> c) & 1) ? d * b[o] : 0.f;
  }
}
void jacobi0(const float* b, float* x, const float* dinv, const uint8_t* mask, float omega, int64_t n, int64_t ld,
             int nrhs, hipStream_t s, const double* skip) {
  hipLaunchKernelGGL(k_jacobi0, dim3((unsigned)((n + 255) / 256), nrhs), dim3(256), 0, s, b, x, dinv, mask, omega, n,
                     ld, skip);
  ANNEAL_HIP(hipGetLastError());
}

// Dense coarse solve x = Ainv b with Ainv in FP64 (n x n, row major).
__global__ void k_dense_solve(const double* __restrict__ Ainv, const float* __restrict__ b, float* __restrict__ x,
                              int nnode, int64_t ld, const OpEntry* __restrict__ ops, int rpo,
                              const double* __restrict__ skip) {
  const int n = 3 * nnode;
  const int r = blockIdx.y;
  if (rhs_skipped(skip, r)) return;
  if (ops) Ainv = ops[r / rpo].Ainv;
  const int row = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
  const int lane = threadIdx.x & 63;
  if (row >= n) return;
  const double* a = Ainv + (int64_t)row * n;
  const float* br = b + (int64_t)3 * r * ld;
  // the row is read in groups of 8 loads per lane so several are in flight;
  // the summation order (j ascending per lane) is the same as one at a time
  double s = 0.0;
  int j = lane;
  for (; j + 7 * 64 < n; j += 8 * 64) {
    double av[8], bv[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int jj = j + u * 64;
      const int node = jj / 3, c = jj - 3 * node;
      av[u] = a[jj];
      bv[u] = (double)br[c * ld + node];
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) s += av[u] * bv[u];
  }
  for (; j < n; j += 64) {
    const int node = j / 3, c = j - 3 * node;
    s += a[j] * (double)br[c * ld + node];
  }
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if (lane == 0) {
    const int node = row / 3, c = row - 3 * node;
    x[(int64_t)(3 * r + c) * ld + node] = (float)s;
  }
}
void dense_solve(const double* Ainv, const float* b, float* x, int nnode, int64_t ld, int nrhs, hipStream_t s,
                 const OpEntry* ops, int rpo, const double* skip) {
  const int n = 3 * nnode;
  hipLaunchKernelGGL(k_dense_solve, dim3((n + 3) / 4, nrhs), dim3(256), 0, s, Ainv, b, x, nnode, ld, ops, rpo, skip);
  ANNEAL_HIP(hipGetLastError());
}

// Dense assembly of the coarsest operator: entry (i, j) gathers the elements
// shared by the two nodes, fixed order.  Non-free dofs get identity rows.
__global__ void k_dense_assemble(const float* __restrict__ Ee, const uint8_t* __restrict__ mask,
                                 const double* __restrict__ ke, double* __restrict__ A, int nx, int ny, int nz) {
  const int nnode = (nx + 1) * (ny + 1) * (nz + 1);
  const int n = 3 * nnode;
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= (int64_t)n * n) return;
  const int i = t / n, j = t % n;
  const int ni = i / 3, ci = i % 3, nj = j / 3, cj = j % 3;
  const bool fi = (mask[ni] >> ci) & 1, fj = (mask[nj] >> cj) & 1;
  double v = 0.0;
  if (!fi || !fj) {
    v = (i == j) ? 1.0 : 0.0;
  } else {
    const int sx = nx + 1, sxy = (nx + 1) * (ny + 1);
    const int ix = ni % sx, iy = (ni / sx) % (ny + 1), iz = ni / sxy;
    const int jx = nj % sx, jy = (nj / sx) % (ny + 1), jz = nj / sxy;
    if (abs(ix - jx) <= 1 && abs(iy - jy) <= 1 && abs(iz - jz) <= 1) {
      for (int ez = max(iz, jz) - 1; ez <= min(iz, jz); ++ez)
        for (int ey = max(iy, jy) - 1; ey <= min(iy, jy); ++ey)
          for (int ex = max(ix, jx) - 1; ex <= min(ix, jx); ++ex) {
            if (ex < 0 || ex >= nx || ey < 0 || ey >= ny || ez < 0 || ez >= nz) continue;
            const double e = Ee[ex + nx * (ey + ny * ez)];
            if (e == 0.0) continue;
            const int qi = (ix - ex) + 2 * (iy - ey) + 4 * (iz - ez);
            const int qj = (jx - ex) + 2 * (jy - ey) + 4 * (jz - ez);
            v += e * ke[(3 * qi + ci) * 24 + 3 * qj + cj];
          }
    }
    if (i == j && v <= 0.0) v = 1.0;  // dof with no stiffness: keep the matrix regular
  }
  A[t] = v;
}

// Block Gauss-Jordan inversion of an SPD matrix (no pivoting), one launch per
// pivot block K of up to kGJB rows, out of place between two buffers:
//   new_KK = P = M_KK^-1, new_KJ = P M_KJ, new_IK = -M_IK P,
//   new_IJ = M_IJ - (M_IK P) M_KJ.
// A workgroup owns kGJB rows x kGJC columns.  Its first wave inverts the small
// pivot block (redundantly in every workgroup); the workgroup then forms
// N = -M_IK P for its rows (N = P on the pivot rows, whose tile starts from 0
// instead of M_IJ), stages the pivot
// rows M_KJ of its columns in LDS, and updates its tile as a register-blocked
// product: each thread owns 4 rows x 8 columns and per pivot index l reads 4
// values of N and 8 of M_KJ for 32 FMAs (a thread per column would read one
// LDS value per FMA, and LDS bandwidth would bound the step).  Blocks are
// zero-padded to kGJB, so the last, partial pivot block runs the same unrolled
// loops.  All global loads are issued first from clamped, in-bounds addresses
// (padding selected afterwards), so they are in flight during the inversion.
constexpr int kGJB = 32, kGJC = 256, kGJR = 4, kGJM = 8;
// Orders the LDS accesses of one wave around this point (no workgroup barrier).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
static_assert(kGJB * kGJC == 256 * kGJR * kGJM, "one 4x8 register block per thread");
__global__ __launch_bounds__(256) void k_bgj_step(const double* __restrict__ A, double* __restrict__ B, int n, int k0,
                                                  int bk) {
  constexpr int kPer = kGJB * kGJB / 256;
  __shared__ double P[kGJB * kGJB];    // pivot block, then its inverse
  __shared__ double Nt[kGJB * kGJB];   // N transposed: Nt[l][row]
  __shared__ double MK[kGJB * kGJC];   // M_KJ; its first kGJB * kGJB entries hold M_IK until N is formed
  double* MI = MK;
  const int tid = threadIdx.x;
  const int r0 = blockIdx.y * kGJB, c0 = blockIdx.x * kGJC;
  const int nr = min(kGJB, n - r0), nc = min(kGJC, n - c0);
  const bool pivot_rows = r0 == k0;
  // loads: pivot block and M_IK entries, this thread's column of M_KJ, its 4x8 tile block
  double pv[kPer], mv[kPer], mk[kGJB], acc[kGJR][kGJM];
#pragma unroll
  for (int q = 0; q < kPer; ++q) {
    const int idx = tid + q * 256, i = idx / kGJB, jj = min(idx % kGJB, bk - 1);
    pv[q] = A[(int64_t)(k0 + min(i, bk - 1)) * n + k0 + jj];
    mv[q] = A[(int64_t)(r0 + min(i, nr - 1)) * n + k0 + jj];
  }
  const int jc = c0 + min(tid, nc - 1);
#pragma unroll
  for (int l = 0; l < kGJB; ++l) mk[l] = A[(int64_t)(k0 + min(l, bk - 1)) * n + jc];
  const int tr = tid / 32, tc = tid % 32;  // rows tr*4 + i, columns tc + 32*m
#pragma unroll
  for (int i = 0; i < kGJR; ++i)
#pragma unroll
    for (int m = 0; m < kGJM; ++m)
      acc[i][m] = A[(int64_t)(r0 + min(tr * kGJR + i, nr - 1)) * n + c0 + min(tc + 32 * m, nc - 1)];
#pragma unroll
  for (int q = 0; q < kPer; ++q) {
    const int idx = tid + q * 256, i = idx / kGJB, jj = idx % kGJB;
    P[idx] = (i < bk && jj < bk) ? pv[q] : 0.0;
    MI[idx] = (i < nr && jj < bk) ? mv[q] : 0.0;
  }
  __syncthreads();
  // Gauss-Jordan on the pivot block by the first wave alone, so the 32
  // serial pivots take no workgroup barrier (the other waves wait at the one
  // after it).  Lane l keeps column pj = l % 32 of rows 16 h + q (h = l / 32,
  // q < 16) in registers; before pivot k the lanes holding row k and column k
  // publish them in LDS (Nt, free until N is formed; double-buffered by pivot
  // parity), and the wave reads them back after a wave-level fence: the LDS
  // accesses of one wave complete in order.
  static_assert(kGJB == 32, "pivot-block entries: one column of 16 rows per lane of one wave");
  if (tid < 64) {
    const int pj = tid % kGJB, h = tid / kGJB;
    double e[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) e[q] = P[(16 * h + q) * kGJB + pj];
    auto publish = [&](int k, double* R) {
      double* C = R + kGJB;
      if (h == (k >> 4)) R[pj] = e[k & 15];
      if (pj == k) {
#pragma unroll
        for (int q = 0; q < 16; ++q) C[16 * h + q] = e[q];
      }
    };
    publish(0, Nt);
#pragma unroll
    for (int k = 0; k < kGJB; ++k) {
      if (k < bk) {
        wave_sync();
        const double* R = Nt + (k & 1) * 2 * kGJB;  // row k, then column k
        const double* C = R + kGJB;
        const double rp = 1.0 / R[k];
        const double akj = R[pj] * rp;
        const bool colk = pj == k;
#pragma unroll
        for (int q = 0; q < 16; ++q) {
          const double aik = C[16 * h + q];
          double v = colk ? -aik * rp : e[q] - aik * akj;
          if (q == (k & 15) && h == (k >> 4)) v = colk ? rp : akj;  // row k
          e[q] = v;
        }
        if (k + 1 < bk) publish(k + 1, Nt + ((k + 1) & 1) * 2 * kGJB);
      }
    }
#pragma unroll
    for (int q = 0; q < 16; ++q) P[(16 * h + q) * kGJB + pj] = e[q];
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < kPer; ++q) {
    const int idx = tid + q * 256, ri = idx / kGJB, pj = idx % kGJB;
    double t = 0.0;
#pragma unroll
    for (int l = 0; l < kGJB; ++l) t += MI[ri * kGJB + l] * P[l * kGJB + pj];
    Nt[pj * kGJB + ri] = pivot_rows ? P[idx] : -t;
  }
  __syncthreads();  // M_IK is consumed: its space takes M_KJ
#pragma unroll
  for (int l = 0; l < kGJB; ++l) MK[l * kGJC + tid] = (tid < nc && l < bk) ? mk[l] : 0.0;
  __syncthreads();
#pragma unroll
  for (int i = 0; i < kGJR; ++i)
#pragma unroll
    for (int m = 0; m < kGJM; ++m)
      if (pivot_rows) acc[i][m] = 0.0;
#pragma unroll 4
  for (int l = 0; l < kGJB; ++l) {
    double nv[kGJR], kv[kGJM];
#pragma unroll
    for (int i = 0; i < kGJR; ++i) nv[i] = Nt[l * kGJB + tr * kGJR + i];
#pragma unroll
    for (int m = 0; m < kGJM; ++m) kv[m] = MK[l * kGJC + tc + 32 * m];
#pragma unroll
    for (int i = 0; i < kGJR; ++i)
#pragma unroll
      for (int m = 0; m < kGJM; ++m) acc[i][m] += nv[i] * kv[m];
  }
#pragma unroll
  for (int i = 0; i < kGJR; ++i) {
    const int ri = tr * kGJR + i;
    if (ri >= nr) continue;
#pragma unroll
    for (int m = 0; m < kGJM; ++m) {
      const int cj = tc + 32 * m, j = c0 + cj;
      if (cj >= nc) continue;
      // pivot columns: -M_IK P, or P itself on the pivot rows
      const double v = (j >= k0 && j < k0 + bk) ? Nt[(j - k0) * kGJB + ri] : acc[i][m];
      B[(int64_t)(r0 + ri) * n + j] = v;
    }
  }
}

void dense_assemble_invert(const float* Ee, const uint8_t* mask, const double* ke, double* A, double* B, int nx,
                           int ny, int nz, hipStream_t s) {
  const int n = 3 * (nx + 1) * (ny + 1) * (nz + 1);
  const int64_t nn = (int64_t)n * n;
  hipLaunchKernelGGL(k_dense_assemble, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, s, Ee, mask, ke, A, nx, ny,
                     nz);
  ANNEAL_HIP(hipGetLastError());
  ANNEAL_CHECK(n <= 3 * 4096, "coarsest multigrid level too large for the dense inverse");
  double* src = A;
  double* dst = B;
  const dim3 grid((unsigned)((n + kGJC - 1) / kGJC), (unsigned)((n + kGJB - 1) / kGJB));
  for (int k0 = 0; k0 < n; k0 += kGJB) {
    const int bk = std::min(kGJB, n - k0);
    hipLaunchKernelGGL(k_bgj_step, grid, dim3(256), 0, s, src, dst, n, k0, bk);
    ANNEAL_HIP(hipGetLastError());
    std::swap(src, dst);
  }
  if (src != A) ANNEAL_HIP(hipMemcpyAsync(A, src, nn * sizeof(double), hipMemcpyDeviceToDevice, s));
}

// ---------------------------------------------------------------------------
// PCG pieces (FP32 vectors, FP64 scalars on the device).
//   st layout per rhs r: st[r*kSt + ...]
// Reduce per-rhs partials (fixed order) straight into the PCG scalar block
// and finish the scalar update: op 0 stores, op 1 stores pq and sets alpha,
// op 2 takes rz_new and sets beta (and rz).
__global__ __launch_bounds__(kRedThreads) void k_finalize_st(const double* __restrict__ part, int nblk,
                                                             double* __restrict__ st, int field, int op) {
  const int r = blockIdx.x;
  double s = 0.0;
  for (int i = threadIdx.x; i < nblk; i += kRedThreads) s += part[(int64_t)r * nblk + i];
  __shared__ double sh[kRedThreads / 64];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x != 0) return;
  double t = 0.0;
  for (int w = 0; w < kRedThreads / 64; ++w) t += sh[w];
  double* q = st + r * kSt;
  q[field] = t;
  const bool live = q[kDone] == 0.0;
  if (op == 1) {
    q[kAlpha] = (live && t > 0.0 && isfinite(t)) ? q[kRZ] / t : 0.0;
  } else if (op == 2) {
    q[kBeta] = (live && q[kRZ] > 0.0) ? t / q[kRZ] : 0.0;
    if (live) q[kRZ] = t;
  } else if (op == 3) {
    const double pq = q[kPQ];
    q[kBeta] = (live && pq > 0.0 && isfinite(t)) ? -t / pq : 0.0;
  } else if (op == 4) {
    if (live) q[kRZ] = t;
  }
}
void finalize_st(const double* part, int nblk, int nrhs, double* st, int field, int op, hipStream_t s) {
  hipLaunchKernelGGL(k_finalize_st, dim3(nrhs), dim3(kRedThreads), 0, s, part, nblk, st, field, op);
  ANNEAL_HIP(hipGetLastError());
}
void dot_partial(const float* a, const float* b, int64_t n, int64_t ld, int nrhs, double* part, hipStream_t s,
                 const double* skip) {
  hipLaunchKernelGGL((k_dot_partial<float, float>), dim3(kDotBlocks, nrhs), dim3(kRedThreads), 0, s, a, b, n, ld,
                     part, skip);
  ANNEAL_HIP(hipGetLastError());
}

// x += alpha p ; r -= alpha q ; partial of r.r
__global__ __launch_bounds__(kRedThreads) void k_pcg_update(float* __restrict__ x, float* __restrict__ rv,
                                                             const float* __restrict__ p, const float* __restrict__ q,
                                                             const double* __restrict__ st, int64_t n, int64_t ld,
                                                             double* __restrict__ part,
                                                             const double* __restrict__ skip) {
  const int r = blockIdx.y;
  if (rhs_skipped(skip, r)) return;
  const float al = (float)st[r * kSt + kAlpha];
  double s = 0.0;
  for (int c = 0; c < 3; ++c) {
    const int64_t o = (int64_t)(3 * r + c) * ld;
    for (int64_t i = blockIdx.x * (int64_t)kRedThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kRedThreads) {
      x[o + i] += al * p[o + i];
      const float rn = rv[o + i] - al * q[o + i];
      rv[o + i] = rn;
      s += (double)rn * (double)rn;
    }
  }
  __shared__ double sh[kRedThreads / 64];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < kRedThreads / 64; ++w) t += sh[w];
    part[(int64_t)r * gridDim.x + blockIdx.x] = t;
  }
}
void pcg_update(float* x, float* r, const float* p, const float* q, const double* st, int64_t n, int64_t ld, int nrhs,
                double* part, double* out_rr, hipStream_t s, const double* skip, double* out_host) {
  hipLaunchKernelGGL(k_pcg_update, dim3(kDotBlocks, nrhs), dim3(kRedThreads), 0, s, x, r, p, q, st, n, ld, part, skip);
  ANNEAL_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_finalize, dim3(nrhs), dim3(kRedThreads), 0, s, part, kDotBlocks, out_rr, out_host);
  ANNEAL_HIP(hipGetLastError());
}

// |r|^2 of the step and the device-side convergence test (one block per rhs;
// the summation order is k_finalize's).  A rhs skipped by the step (done
// before it) left no partials and is not tested again.
__global__ __launch_bounds__(kRedThreads) void k_finalize_converge(const double* __restrict__ part, int nblk,
                                                                   double* __restrict__ st,
                                                                   double* __restrict__ rr_host,
                                                                   double* __restrict__ done_host) {
  const int r = blockIdx.x;
  double* q = st + r * kSt;
  const bool live = q[kDone] == 0.0;
  double s = 0.0;
  for (int i = threadIdx.x; i < nblk; i += kRedThreads) s += part[(int64_t)r * nblk + i];
  __shared__ double sh[kRedThreads / 64];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x != 0) return;
  double t = 0.0;
  for (int w = 0; w < kRedThreads / 64; ++w) t += sh[w];
  if (live) {
    rr_host[r] = t;
    if (t <= q[kTol2]) q[kDone] = 1.0;
  }
  done_host[r] = q[kDone];
}
void pcg_update_converge(float* x, float* r, const float* p, const float* q, double* st, int64_t n, int64_t ld,
                         int nrhs, double* part, hipStream_t s, double* rr_host, double* done_host) {
  hipLaunchKernelGGL(k_pcg_update, dim3(kDotBlocks, nrhs), dim3(kRedThreads), 0, s, x, r, p, q, st, n, ld, part,
                     (const double*)st);
  ANNEAL_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_finalize_converge, dim3(nrhs), dim3(kRedThreads), 0, s, part, kDotBlocks, st, rr_host,
                     done_host);
  ANNEAL_HIP(hipGetLastError());
}

// p = z + beta p
__global__ void k_pcg_dir(float* __restrict__ p, const float* __restrict__ z, const double* __restrict__ st, int64_t n,
                          int64_t ld, int first) {
  const int r = blockIdx.y;
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float be = first ? 0.f : (float)st[r * kSt + kBeta];
  const bool done = st[r * kSt + kDone] != 0.0;
  if (done && !first) return;
  for (int c = 0; c < 3; ++c) {
    const int64_t o = (int64_t)(3 * r + c) * ld + i;
    p[o] = z[o] + be * p[o];
  }
}
void pcg_dir(float* p, const float* z, const double* st, int64_t n, int64_t ld, int nrhs, int first, hipStream_t s) {
  hipLaunchKernelGGL(k_pcg_dir, dim3((unsigned)((n + 255) / 256), nrhs), dim3(256), 0, s, p, z, st, n, ld, first);
  ANNEAL_HIP(hipGetLastError());
}

__global__ __launch_bounds__(256) void k_upload(const double* __restrict__ src1, double* __restrict__ dst1, int n1,
                                                const double* __restrict__ src2, double* __restrict__ dst2, int n2) {
  for (int i = threadIdx.x; i < n1; i += 256) dst1[i] = src1[i];
  for (int i = threadIdx.x; i < n2; i += 256) dst2[i] = src2[i];
}
__global__ __launch_bounds__(64) void k_pass_setup(const double* __restrict__ rr, const double* __restrict__ ff,
                                                   double rtol, double factor, double floor_,
                                                   double* __restrict__ scale, double* __restrict__ st, int nrhs) {
  const int r = threadIdx.x;
  if (r >= nrhs) return;
  const double rn = sqrt(rr[r]), fn = sqrt(ff[r]);
  const bool done = rn <= rtol * fn;
  scale[r] = done ? 0.0 : 1.0 / rn;
  for (int k = 0; k < kSt; ++k) st[r * kSt + k] = 0.0;
  st[r * kSt + kDone] = done ? 1.0 : 0.0;
  const double t = fmax(done ? 1.0 : factor * rtol * fn / rn, floor_);
  st[r * kSt + kTol2] = t * t;
}
void pass_setup(const double* rr, const double* ff, double rtol, double factor, double floor_, double* scale,
                double* st, int nrhs, hipStream_t s) {
  hipLaunchKernelGGL(k_pass_setup, dim3(1), dim3(64), 0, s, rr, ff, rtol, factor, floor_, scale, st, nrhs);
  ANNEAL_HIP(hipGetLastError());
}

void upload(const double* src1, double* dst1, int n1, const double* src2, double* dst2, int n2, hipStream_t s) {
  hipLaunchKernelGGL(k_upload, dim3(1), dim3(256), 0, s, src1, dst1, n1, src2, dst2, n2);
  ANNEAL_HIP(hipGetLastError());
}

// Refinement glue: r32 = scale_r * r64 ; x64 += inv_scale_r * x32
__global__ void k_to_f32_scaled(const double* __restrict__ a, float* __restrict__ b, const double* __restrict__ sc,
                                int64_t n, int64_t ld) {
  const int r = blockIdx.y;
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double s = sc[r];
  for (int c = 0; c < 3; ++c) {
    const int64_t o = (int64_t)(3 * r + c) * ld + i;
    b[o] = (float)(a[o] * s);
  }
}
__global__ void k_axpy_f32_to_f64(const float* __restrict__ x, double* __restrict__ y, const double* __restrict__ sc,
                                  int64_t n, int64_t ld) {
  const int r = blockIdx.y;
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double s = sc[r];
  for (int c = 0; c < 3; ++c) {
    const int64_t o = (int64_t)(3 * r + c) * ld + i;
    y[o] += s * (double)x[o];
  }
}
void to_f32_scaled(const double* a, float* b, const double* sc, int64_t n, int64_t ld, int nrhs, hipStream_t s) {
  hipLaunchKernelGGL(k_to_f32_scaled, dim3((unsigned)((n + 255) / 256), nrhs), dim3(256), 0, s, a, b, sc, n, ld);
  ANNEAL_HIP(hipGetLastError());
}
void axpy_f32_to_f64(const float* x, double* y, const double* sc, int64_t n, int64_t ld, int nrhs, hipStream_t s) {
  hipLaunchKernelGGL(k_axpy_f32_to_f64, dim3((unsigned)((n + 255) / 256), nrhs), dim3(256), 0, s, x, y, sc, n, ld);
  ANNEAL_HIP(hipGetLastError());
}

__global__ void k_cast_f64_f32(const double* __restrict__ a, float* __restrict__ b, int64_t n) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) b[i] = (float)a[i];
}
void cast_f64_f32(const double* a, float* b, int64_t n, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_cast_f64_f32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a, b, n);
  ANNEAL_HIP(hipGetLastError());
}

__global__ void k_copy_cast_f64(const double* __restrict__ a, double* __restrict__ a64, float* __restrict__ b,
                                int64_t n) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double v = a[i];
  a64[i] = v;
  b[i] = (float)v;
}
void copy_cast_f64(const double* a, double* a64, float* b, int64_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_copy_cast_f64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a, a64, b, n);
  ANNEAL_HIP(hipGetLastError());
}

}  // namespace anneal
