// Glue between the multigrid hierarchy and the U-Net: network input/output on
// the node grid of one coarse level.
//
// Input channels per node (bf16): 0 density feature cbrt(mean of the adjacent
// elements' Ee / e_ref); 1-3 the level's residual scaled to unit RMS; 4-6 the
// fixed-dof flags (1 where component c is not free); 7 the active-node flag
// (some adjacent element has Ee > 0).  Output (FP32, 3 channels) is the
// displacement correction in the same units: z = out * |r| / sqrt(V) / e_ref,
// so the map is homogeneous of degree one in r.
#include <hip/hip_runtime.h>

#include <cmath>

#include "common.hpp"
#include "learned.hpp"

namespace anneal {
namespace {

__device__ __forceinline__ uint16_t f2bf(float f) { return __builtin_bit_cast(uint16_t, (__bf16)f); }

__device__ __forceinline__ float node_density(const float* Ee, int nx, int ny, int nz, int i, int j, int k,
                                              float inv_eref, bool& active) {
  float s = 0.f;
  int n = 0;
  active = false;
  for (int c = 0; c < 8; ++c) {
    const int ei = i - 1 + (c & 1), ej = j - 1 + ((c >> 1) & 1), ek = k - 1 + (c >> 2);
    if (ei < 0 || ej < 0 || ek < 0 || ei >= nx || ej >= ny || ek >= nz) continue;
    const float e = Ee[ei + (int64_t)nx * (ej + (int64_t)ny * ek)];
    s += e;
    ++n;
    active = active || e > 0.f;
  }
  return n ? cbrtf(fmaxf(s / n, 0.f) * inv_eref) : 0.f;
}

__global__ void k_pack(const float* __restrict__ Ee, const uint8_t* __restrict__ mask, const float* __restrict__ r,
                       const double* __restrict__ nrm2, int nx, int ny, int nz, int64_t ld, float inv_eref,
                       uint16_t* __restrict__ out) {
  const int64_t nn = (int64_t)(nx + 1) * (ny + 1) * (nz + 1);
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  if (n >= nn) return;
  const int i = int(n % (nx + 1)), j = int((n / (nx + 1)) % (ny + 1)), k = int(n / ((int64_t)(nx + 1) * (ny + 1)));
  bool act;
  const float rho = node_density(Ee, nx, ny, nz, i, j, k, inv_eref, act);
  const double s2 = nrm2[b];
  const float sc = s2 > 0.0 ? float(sqrt(double(nn) / s2)) : 0.f;
  const uint8_t m = mask[n];
  uint16_t v[8];
  v[0] = f2bf(rho);
  for (int c = 0; c < 3; ++c) v[1 + c] = f2bf(sc * r[(int64_t)(3 * b + c) * ld + n]);
  for (int c = 0; c < 3; ++c) v[4 + c] = f2bf(((m >> c) & 1) ? 0.f : 1.f);
  v[7] = f2bf(act ? 1.f : 0.f);
  uint16_t* o = out + ((int64_t)b * nn + n) * 8;
  for (int c = 0; c < 8; ++c) o[c] = v[c];
}

__global__ void k_unpack(const float* __restrict__ o, const uint8_t* __restrict__ mask, const double* __restrict__ nrm2,
                         int64_t nn, int64_t ld, float inv_eref, float* __restrict__ z) {
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  if (n >= ld) return;
  float sc = 0.f;
  if (n < nn) sc = float(sqrt(nrm2[b] / double(nn))) * inv_eref;
  const uint8_t m = n < nn ? mask[n] : 0;
  for (int c = 0; c < 3; ++c)
    z[(int64_t)(3 * b + c) * ld + n] = ((m >> c) & 1) ? sc * o[((int64_t)b * nn + n) * 3 + c] : 0.f;
}

// target[b][n][c] = U[b][c][fine node (2^l i, 2^l j, 2^l k)] * sqrt(V) / |r| * e_ref
__global__ void k_target(const double* __restrict__ U, int64_t ldf, int fnx, int fny, int step, int nx, int ny, int nz,
                         const double* __restrict__ nrm2, float eref, float* __restrict__ t) {
  const int64_t nn = (int64_t)(nx + 1) * (ny + 1) * (nz + 1);
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  if (n >= nn) return;
  const int i = int(n % (nx + 1)), j = int((n / (nx + 1)) % (ny + 1)), k = int(n / ((int64_t)(nx + 1) * (ny + 1)));
  const int64_t fn = (int64_t)step * i + (int64_t)(fnx + 1) * ((int64_t)step * j + (int64_t)(fny + 1) * step * k);
  const double s2 = nrm2[b];
  const double sc = s2 > 0.0 ? sqrt(double(nn) / s2) * eref : 0.0;
  for (int c = 0; c < 3; ++c) t[((int64_t)b * nn + n) * 3 + c] = float(sc * U[(int64_t)(3 * b + c) * ldf + fn]);
}

__global__ void k_zero(float* __restrict__ x, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = 0.f;
}

__global__ void k_add(float* __restrict__ y, const float* __restrict__ x, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] += x[i];
}

// gamma_b = <r, z1> / <z1, K z1> (0 when the denominator is not positive);
// t = r - gamma K z1 and the caller's z1 scaled by gamma.
__global__ void k_gamma(const double* __restrict__ rz, const double* __restrict__ zkz, int nrhs,
                        double* __restrict__ gamma) {
  const int b = threadIdx.x;
  if (b >= nrhs) return;
  const double d = zkz[b];
  gamma[b] = (d > 0.0 && isfinite(d) && isfinite(rz[b])) ? rz[b] / d : 0.0;
}

__global__ void k_scaled_resid(const float* __restrict__ r, const float* __restrict__ kz, float* __restrict__ z1,
                               float* __restrict__ t, const double* __restrict__ gamma, int64_t ld) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  if (i >= 3 * ld) return;
  const float g = float(gamma[b]);
  const int64_t o = (int64_t)b * 3 * ld + i;
  t[o] = r[o] - g * kz[o];
  z1[o] = g * z1[o];
}

inline unsigned nb(int64_t n) { return unsigned((n + 255) / 256); }

}  // namespace

void net_pack(const float* Ee, const uint8_t* mask, const float* r, const double* nrm2, const GridDims& g, int nrhs,
              double e_ref, uint16_t* out, hipStream_t s) {
  hipLaunchKernelGGL(k_pack, dim3(nb(g.nnode()), nrhs), dim3(256), 0, s, Ee, mask, r, nrm2, g.nx, g.ny, g.nz, g.ldn(),
                     float(1.0 / e_ref), out);
  ANNEAL_HIP(hipGetLastError());
}

void net_unpack(const float* o, const uint8_t* mask, const double* nrm2, const GridDims& g, int nrhs, double e_ref,
                float* z, hipStream_t s) {
  hipLaunchKernelGGL(k_unpack, dim3(nb(g.ldn()), nrhs), dim3(256), 0, s, o, mask, nrm2, g.nnode(), g.ldn(),
                     float(1.0 / e_ref), z);
  ANNEAL_HIP(hipGetLastError());
}

void net_target(const double* U, const GridDims& fine, int level, const GridDims& g, const double* nrm2, int nrhs,
                double e_ref, float* t, hipStream_t s) {
  hipLaunchKernelGGL(k_target, dim3(nb(g.nnode()), nrhs), dim3(256), 0, s, U, fine.ldn(), fine.nx, fine.ny, 1 << level,
                     g.nx, g.ny, g.nz, nrm2, float(e_ref), t);
  ANNEAL_HIP(hipGetLastError());
}

void net_gamma(const double* rz, const double* zkz, int nrhs, double* gamma, hipStream_t s) {
  hipLaunchKernelGGL(k_gamma, dim3(1), dim3(64), 0, s, rz, zkz, nrhs, gamma);
  ANNEAL_HIP(hipGetLastError());
}

void net_scaled_resid(const float* r, const float* kz, float* z1, float* t, const double* gamma, int64_t ld, int nrhs,
                      hipStream_t s) {
  hipLaunchKernelGGL(k_scaled_resid, dim3(nb(3 * ld), nrhs), dim3(256), 0, s, r, kz, z1, t, gamma, ld);
  ANNEAL_HIP(hipGetLastError());
}

void fill_zero(float* x, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_zero, dim3(nb(n)), dim3(256), 0, s, x, n);
  ANNEAL_HIP(hipGetLastError());
}

void add_f32(float* y, const float* x, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_add, dim3(nb(n)), dim3(256), 0, s, y, x, n);
  ANNEAL_HIP(hipGetLastError());
}

}  // namespace anneal
