// Display-mesh sampling of saved result fields (anneal/fields.py:17-37).
//
// One thread per triangle corner:
//   * displacement: trilinear interpolation of the node field u[nnode][3]
//     (float32, node index i + (nx+1)(j + (ny+1)k)) at (p - o) / h, with the
//     coordinate clamped to the node box (scipy map_coordinates, order 1,
//     mode "nearest", fields.py:26-28);
//   * stress: the element value vm[nelem] after a 2x2x2 grey dilation
//     (window [i, i+1] per axis, reflected at the upper face, fields.py:30-31),
//     looked up with order 0 at (p - 0.6 h n - o) / h - 0.5 (fields.py:32-35).
// The dilation is folded into the lookup: the corner reads the 8 voxels of its
// window directly, so no dilated copy of vm is stored.
//
// Arithmetic follows the library's evaluation order so the float32 result is
// the same rounding of the same FP64 value: per axis t = x - floor(x),
// w0 = 1 - t, w1 = 1 - w0; each corner term is ((v * wx) * wy) * wz, summed
// with z fastest.  Contraction into FMAs is switched off for that reason.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "common.hpp"
#include "fields.hpp"

namespace anneal {
namespace {

struct Axis {
  int64_t i0, i1;
  double w0, w1;
};

__device__ __forceinline__ Axis linear_axis(double x, int64_t len) {
#pragma clang fp contract(off)
  if (x < 0.0) x = 0.0;
  else if (x > double(len - 1)) x = double(len - 1);
  const double f = floor(x);
  Axis a;
  a.i0 = int64_t(f);
  a.i1 = a.i0 + 1 < len ? a.i0 + 1 : len - 1;
  const double t = x - f;
  a.w0 = 1.0 - t;
  a.w1 = 1.0 - a.w0;
  return a;
}

__device__ __forceinline__ int64_t nearest_index(double x, int64_t len) {
  if (x < 0.0) x = 0.0;
  else if (x > double(len - 1)) x = double(len - 1);
  return int64_t(floor(x + 0.5));
}

__global__ void __launch_bounds__(256) k_corner_values(const float* __restrict__ u, const float* __restrict__ vm,
                                                       int nx, int ny, int nz, double h, double ox, double oy,
                                                       double oz, const double* __restrict__ tri,
                                                       const double* __restrict__ nrm, int64_t ncorner,
                                                       float* __restrict__ out) {
#pragma clang fp contract(off)
  const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= ncorner) return;
  const double px = tri[3 * t + 0], py = tri[3 * t + 1], pz = tri[3 * t + 2];
  const int64_t f = t / 3;
  const double nxv = nrm[3 * f + 0], nyv = nrm[3 * f + 1], nzv = nrm[3 * f + 2];
  const int64_t sx = nx + 1, sxy = int64_t(nx + 1) * (ny + 1);

  // displacement
  const Axis ax = linear_axis((px - ox) / h, nx + 1);
  const Axis ay = linear_axis((py - oy) / h, ny + 1);
  const Axis az = linear_axis((pz - oz) / h, nz + 1);
  double acc[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int a = 0; a < 2; ++a) {
    const int64_t i = a ? ax.i1 : ax.i0;
    const double wx = a ? ax.w1 : ax.w0;
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const int64_t j = b ? ay.i1 : ay.i0;
      const double wy = b ? ay.w1 : ay.w0;
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        const int64_t k = c ? az.i1 : az.i0;
        const double wz = c ? az.w1 : az.w0;
        const int64_t n = i + sx * j + sxy * k;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          double v = double(u[3 * n + d]);
          v = v * wx;
          v = v * wy;
          v = v * wz;
          acc[d] = acc[d] + v;
        }
      }
    }
  }

  // von Mises of the dilated element field just inside the surface
  const double s = 0.6 * h;
  const double qx = px - s * nxv, qy = py - s * nyv, qz = pz - s * nzv;
  const int64_t ei = nearest_index((qx - ox) / h - 0.5, nx);
  const int64_t ej = nearest_index((qy - oy) / h - 0.5, ny);
  const int64_t ek = nearest_index((qz - oz) / h - 0.5, nz);
  float m = -INFINITY;
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    const int64_t k = ek + c < nz ? ek + c : nz - 1;
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const int64_t j = ej + b < ny ? ej + b : ny - 1;
#pragma unroll
      for (int a = 0; a < 2; ++a) {
        const int64_t i = ei + a < nx ? ei + a : nx - 1;
        m = fmaxf(m, vm[i + int64_t(nx) * (j + int64_t(ny) * k)]);
      }
    }
  }
  float4 o;
  o.x = float(acc[0]);
  o.y = float(acc[1]);
  o.z = float(acc[2]);
  o.w = m;
  reinterpret_cast<float4*>(out)[t] = o;
}

}  // namespace

void corner_values(const float* u, const float* vm, int nx, int ny, int nz, double h, const double* origin,
                   const double* tri, const double* nrm, int64_t nface, float* out, hipStream_t s) {
  const int64_t nc = 3 * nface;
  if (nc == 0) return;
  const int64_t blocks = (nc + 255) / 256;
  ANNEAL_CHECK(blocks < (int64_t(1) << 31), "too many triangle corners");
  hipLaunchKernelGGL(k_corner_values, dim3(unsigned(blocks)), dim3(256), 0, s, u, vm, nx, ny, nz, h, origin[0],
                     origin[1], origin[2], tri, nrm, nc, out);
  ANNEAL_HIP(hipGetLastError());
}

}  // namespace anneal
