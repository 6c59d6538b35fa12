// 3-D U-Net on node grids with BF16 MFMA implicit-GEMM convolutions (gfx950).
//
// Convolution as a GEMM: M = voxels, N = output channels, K = taps x input
// channels (k = tap * cin + ci).  One wave computes a (16 MB) x (16 NB) output
// tile with v_mfma_f32_16x16x32_bf16: lane l supplies row l%16 of A (its
// voxel's 8 consecutive k = one 16-byte load of 8 channels at one tap) and
// column l%16 of B (8 consecutive k of one output channel's weight row), and
// receives rows 4(l/16)..+3 of column l%16 of the FP32 result.  The data
// gradient is the same kernel on the output gradient with the taps reversed and
// the channel roles swapped; the weight gradient reduces over voxels, so its
// operands are transposed through LDS (voxels become the contiguous K index).
//
// Every reduction has a fixed order (fixed voxel chunks, then a sequential sum
// over chunks), so training and inference are bit-reproducible.
#include "unet.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>

namespace anneal {

struct ConvK {
  const uint16_t* x;
  int xpitch, cin, taps;
  const uint16_t* w;  // [n][kp]
  int kp, nchunk;
  const float* bias;  // [n] or null
  uint16_t* y;        // bf16 output view or null
  int ypitch;
  float* yf;  // FP32 output [v][cf] (channels < cf) or null
  int cf;
  const uint16_t* mask;  // zero the output where mask <= 0 (or null)
  int mpitch;
  int relu, accum;
  int nx, ny, nz;
  int64_t nvox;
  // split-K: blockIdx.z covers chunks [z * kper, (z + 1) * kper); with
  // ksplit > 1 the raw sums go to part[z][v][n] and k_conv_finish applies the
  // epilogue (fixed summation order over z)
  int ksplit, kper, nout;
  float* part;
};

namespace {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));
typedef unsigned u32x4_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float(uint32_t(h) << 16); }
__device__ __forceinline__ uint16_t f2bf(float f) { return __builtin_bit_cast(uint16_t, (__bf16)f); }
__device__ __forceinline__ u32x4_t ld16(const uint16_t* p) { return *reinterpret_cast<const u32x4_t*>(p); }
__device__ __forceinline__ void unpack8(u32x4_t u, float* f) {
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    f[2 * i] = __uint_as_float(u[i] << 16);
    f[2 * i + 1] = __uint_as_float(u[i] & 0xffff0000u);
  }
}
__device__ __forceinline__ u32x4_t pack8(const float* f) {
  u32x4_t u;
#pragma unroll
  for (int i = 0; i < 4; ++i) u[i] = uint32_t(f2bf(f[2 * i])) | (uint32_t(f2bf(f[2 * i + 1])) << 16);
  return u;
}

__device__ __forceinline__ void tap_offsets(int tap, int taps, int& dx, int& dy, int& dz) {
  if (taps == 27) {
    dz = tap / 9;
    const int t2 = tap - 9 * dz;
    dy = t2 / 3;
    dx = t2 - 3 * dy - 1;
    dy -= 1;
    dz -= 1;
  } else {
    dx = dy = dz = 0;
  }
}


template <int NB, int MB>
__global__ void __launch_bounds__(256) k_conv(ConvK a) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, g = lane >> 4, r = lane & 15;
  const int n0 = blockIdx.y * (16 * NB);
  const int64_t vb = ((int64_t)blockIdx.x * 4 + wv) * (16 * MB);
  const int64_t nxy = (int64_t)a.nx * a.ny, nvs = nxy * a.nz;
  int64_t vrow[MB];
  int ci[MB], cj[MB], ck[MB];
  bool vin[MB];
#pragma unroll
  for (int mb = 0; mb < MB; ++mb) {
    const int64_t v = vb + 16 * mb + r;
    vin[mb] = v < a.nvox;
    const int64_t loc = vin[mb] ? v % nvs : 0;
    ck[mb] = int(loc / nxy);
    const int64_t rem = loc - (int64_t)ck[mb] * nxy;
    cj[mb] = int(rem / a.nx);
    ci[mb] = int(rem - (int64_t)cj[mb] * a.nx);
    vrow[mb] = v;
  }
  f32x4_t acc[MB][NB];
#pragma unroll
  for (int mb = 0; mb < MB; ++mb)
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) acc[mb][nb] = f32x4_t{0.f, 0.f, 0.f, 0.f};

  // this lane group's k = 8 g + 32 c  ->  (tap, c0), advanced by 32 per chunk
  const int cbeg = blockIdx.z * a.kper, cend = min(a.nchunk, cbeg + a.kper);
  int tap = (32 * cbeg + 8 * g) / a.cin, c0 = (32 * cbeg + 8 * g) % a.cin;
  int dx, dy, dz;
  tap_offsets(tap, a.taps, dx, dy, dz);
  // fragments of chunk c are loaded one chunk ahead of their MFMAs
  auto load = [&](int c, bf16x8_t (&A)[MB], bf16x8_t (&Bv)[NB]) {
    const int k0 = 32 * c + 8 * g;
    const bool kv = tap < a.taps;
    const int64_t off = dx + (int64_t)dy * a.nx + (int64_t)dz * nxy;
#pragma unroll
    for (int mb = 0; mb < MB; ++mb) {
      const bool ok = kv && vin[mb] && unsigned(ci[mb] + dx) < unsigned(a.nx) && unsigned(cj[mb] + dy) < unsigned(a.ny) &&
                      unsigned(ck[mb] + dz) < unsigned(a.nz);
      u32x4_t u = {0u, 0u, 0u, 0u};
      if (ok) u = ld16(a.x + (vrow[mb] + off) * a.xpitch + c0);
      A[mb] = __builtin_bit_cast(bf16x8_t, u);
    }
#pragma unroll
    for (int nb = 0; nb < NB; ++nb)
      Bv[nb] = __builtin_bit_cast(bf16x8_t, ld16(a.w + (int64_t)(n0 + 16 * nb + r) * a.kp + k0));
    c0 += 32;
    if (c0 >= a.cin) {
      while (c0 >= a.cin) {
        c0 -= a.cin;
        ++tap;
      }
      tap_offsets(tap, a.taps, dx, dy, dz);
    }
  };
  bf16x8_t A0[MB], B0[NB], A1[MB], B1[NB];
  if (cbeg < cend) load(cbeg, A0, B0);
  for (int c = cbeg; c < cend; ++c) {
    if (c + 1 < cend) load(c + 1, A1, B1);
#pragma unroll
    for (int mb = 0; mb < MB; ++mb)
#pragma unroll
      for (int nb = 0; nb < NB; ++nb) acc[mb][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A0[mb], B0[nb], acc[mb][nb], 0, 0, 0);
#pragma unroll
    for (int mb = 0; mb < MB; ++mb) A0[mb] = A1[mb];
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) B0[nb] = B1[nb];
  }
  if (a.ksplit > 1) {
#pragma unroll
    for (int mb = 0; mb < MB; ++mb)
#pragma unroll
      for (int nb = 0; nb < NB; ++nb)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int64_t v = vb + 16 * mb + 4 * g + i;
          if (v < a.nvox) a.part[((int64_t)blockIdx.z * a.nvox + v) * a.nout + n0 + 16 * nb + r] = acc[mb][nb][i];
        }
    return;
  }
#pragma unroll
  for (int mb = 0; mb < MB; ++mb)
#pragma unroll
    for (int nb = 0; nb < NB; ++nb)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int64_t v = vb + 16 * mb + 4 * g + i;
        const int co = n0 + 16 * nb + r;
        if (v >= a.nvox) continue;
        float val = acc[mb][nb][i];
        if (a.bias) val += a.bias[co];
        if (a.accum) val += bf2f(a.y[v * a.ypitch + co]);
        if (a.relu) val = fmaxf(val, 0.f);
        if (a.mask && !(bf2f(a.mask[v * a.mpitch + co]) > 0.f)) val = 0.f;
        if (a.y) a.y[v * a.ypitch + co] = f2bf(val);
        if (a.yf && co < a.cf) a.yf[v * a.cf + co] = val;
      }
}

__global__ void __launch_bounds__(256) k_conv_finish(ConvK a) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.nvox * a.nout) return;
  const int64_t v = t / a.nout;
  const int co = int(t - v * a.nout);
  float val = 0.f;
  for (int z = 0; z < a.ksplit; ++z) val += a.part[((int64_t)z * a.nvox + v) * a.nout + co];
  if (a.bias) val += a.bias[co];
  if (a.accum) val += bf2f(a.y[v * a.ypitch + co]);
  if (a.relu) val = fmaxf(val, 0.f);
  if (a.mask && !(bf2f(a.mask[v * a.mpitch + co]) > 0.f)) val = 0.f;
  if (a.y) a.y[v * a.ypitch + co] = f2bf(val);
  if (a.yf && co < a.cf) a.yf[v * a.cf + co] = val;
}

// 3x3x3 convolution with the input brick staged in LDS (grids with enough
// bricks to fill the machine).  A workgroup owns a brick of 8 (x) x 8 (y) x MB
// (z) output voxels of one sample: MFMA row blocks of 16 voxels (8 x by 2 y);
// wave w takes the y pair w of every z layer (MB row blocks), times NB blocks
// of 16 output channels.  The input brick with its one-voxel halo (10 x 10 x
// (MB + 2) voxels) is staged for S input channels at a time (S = 32, 16 or 8,
// a divisor of cin), so each activation is read from global memory once per
// workgroup and the 27 taps read it from LDS.  Within a slice the K index runs
// over (tap, channel): chunk c, lane group g covers kk = 32 c + 8 g -> tap
// kk / S, channels kk % S .. +7; the chunk loop is unrolled, so the tap
// offsets are constants (taps past 26 pad the last chunk with zero weights).
// The summation order is fixed (slices, then chunks), so results are
// reproducible.
constexpr int kBX = 8, kBY = 8, kHX = kBX + 2, kHY = kBY + 2;

__host__ __device__ constexpr int brick_tap_offset(int tap) {
  return (tap % 3 - 1) + kHX * ((tap / 3) % 3 - 1) + kHX * kHY * (tap / 9 - 1);
}

template <int NB, int S, int MB>
__global__ void __launch_bounds__(256) k_conv_brick(ConvK a) {
  constexpr int PITCH = S + 8;  // bf16 per staged voxel: 16-byte aligned rows, spread over the banks
  constexpr int HV = kHX * kHY * (MB + 2);
  constexpr int NCH = (27 * S + 31) / 32;  // chunks per slice
  __shared__ __attribute__((aligned(16))) uint16_t xs[HV * PITCH];
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6, g = lane >> 4, r = lane & 15;
  const int nbx = (a.nx + kBX - 1) / kBX, nby = (a.ny + kBY - 1) / kBY, nbz = (a.nz + MB - 1) / MB;
  int bid = blockIdx.x;
  const int bx = bid % nbx;
  bid /= nbx;
  const int by = bid % nby;
  bid /= nby;
  const int bz = bid % nbz;
  const int b = bid / nbz;
  const int X0 = bx * kBX, Y0 = by * kBY, Z0 = bz * MB;
  const int64_t nxy = (int64_t)a.nx * a.ny, nvs = nxy * a.nz, vs0 = (int64_t)b * nvs;
  // LDS voxel of this lane's output in row block mb (z layer mb, y pair wv), at tap offset 0
  const int hv0 = ((r & 7) + 1) + kHX * ((2 * wv + (r >> 3)) + 1) + kHX * kHY * 1;
  f32x4_t acc[MB][NB];
#pragma unroll
  for (int mb = 0; mb < MB; ++mb)
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) acc[mb][nb] = f32x4_t{0.f, 0.f, 0.f, 0.f};
  const int n0 = blockIdx.y * 16 * NB;
  const uint16_t* wrow[NB];
#pragma unroll
  for (int nb = 0; nb < NB; ++nb) wrow[nb] = a.w + (int64_t)(n0 + 16 * nb + r) * a.kp;
  for (int s0 = 0; s0 < a.cin; s0 += S) {
    // stage the halo brick for channels s0 .. s0 + S - 1 (zeros outside the grid)
    __syncthreads();  // the previous slice's reads are done
    for (int q = t; q < HV * (S / 8); q += 256) {
      const int hv = q / (S / 8), cg = q % (S / 8);
      const int hx = hv % kHX, hy = (hv / kHX) % kHY, hz = hv / (kHX * kHY);
      const int gx = X0 - 1 + hx, gy = Y0 - 1 + hy, gz = Z0 - 1 + hz;
      u32x4_t u = {0u, 0u, 0u, 0u};
      if (unsigned(gx) < unsigned(a.nx) && unsigned(gy) < unsigned(a.ny) && unsigned(gz) < unsigned(a.nz))
        u = ld16(a.x + (vs0 + gx + (int64_t)a.nx * gy + nxy * gz) * a.xpitch + s0 + 8 * cg);
      *reinterpret_cast<u32x4_t*>(&xs[hv * PITCH + 8 * cg]) = u;
    }
    __syncthreads();
    // chunk c: tap = (32 c + 8 g) / S (lane-group dependent unless S = 32), channels cc = (8 g) % S
    auto frag = [&](int c, bf16x8_t (&A)[MB], bf16x8_t (&Bv)[NB]) {
      const int kk = 32 * c + 8 * g;
      const int tap = kk / S, cc = kk % S;
      const bool kv = tap < 27;
      const int ta = kv ? tap : 26;
      const int off = (hv0 + brick_tap_offset(ta)) * PITCH + cc;
#pragma unroll
      for (int mb = 0; mb < MB; ++mb)
        A[mb] = __builtin_bit_cast(bf16x8_t, *reinterpret_cast<const u32x4_t*>(&xs[off + mb * kHX * kHY * PITCH]));
      const int kw = ta * a.cin + s0 + cc;
#pragma unroll
      for (int nb = 0; nb < NB; ++nb) {
        u32x4_t u = ld16(wrow[nb] + kw);
        if (!kv) u = u32x4_t{0u, 0u, 0u, 0u};
        Bv[nb] = __builtin_bit_cast(bf16x8_t, u);
      }
    };
    bf16x8_t A0[MB], B0[NB], A1[MB], B1[NB];
    frag(0, A0, B0);
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
      if (c + 1 < NCH) frag(c + 1, A1, B1);
#pragma unroll
      for (int mb = 0; mb < MB; ++mb)
#pragma unroll
        for (int nb = 0; nb < NB; ++nb)
          acc[mb][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A0[mb], B0[nb], acc[mb][nb], 0, 0, 0);
#pragma unroll
      for (int mb = 0; mb < MB; ++mb) A0[mb] = A1[mb];
#pragma unroll
      for (int nb = 0; nb < NB; ++nb) B0[nb] = B1[nb];
    }
  }
  // epilogue: lane holds rows 4 g .. 4 g + 3 of its row block, column r of each N block
#pragma unroll
  for (int mb = 0; mb < MB; ++mb)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int row = 4 * g + i;
      const int x = X0 + (row & 7), y = Y0 + 2 * wv + (row >> 3), z = Z0 + mb;
      if (x >= a.nx || y >= a.ny || z >= a.nz) continue;
      const int64_t v = vs0 + x + (int64_t)a.nx * y + nxy * z;
#pragma unroll
      for (int nb = 0; nb < NB; ++nb) {
        const int co = n0 + 16 * nb + r;
        float val = acc[mb][nb][i];
        if (a.bias) val += a.bias[co];
        if (a.accum) val += bf2f(a.y[v * a.ypitch + co]);
        if (a.relu) val = fmaxf(val, 0.f);
        if (a.mask && !(bf2f(a.mask[v * a.mpitch + co]) > 0.f)) val = 0.f;
        if (a.y) a.y[v * a.ypitch + co] = f2bf(val);
        if (a.yf && co < a.cf) a.yf[v * a.cf + co] = val;
      }
    }
}

struct WgradK {
  const uint16_t* x;
  int xpitch, cin, taps;
  const uint16_t* dy;
  int dypitch, cout;  // cout: padded output channels (16 NBW)
  float* part;        // [nchunk][cout][kp]
  float* bpart;       // [nchunk][cout]
  int kp;
  int nx, ny, nz;
  int64_t nvox;
  int steps;  // 32-voxel steps per chunk
};

constexpr int kLdsPitch = 40;  // 32 voxels + 8 pad (bf16), rows stay 16-byte aligned

template <int NBW>
__global__ void __launch_bounds__(256) k_wgrad(WgradK a) {
  __shared__ __attribute__((aligned(16))) uint16_t sdy[64 * kLdsPitch];  // [co][v]
  __shared__ __attribute__((aligned(16))) uint16_t sx[64 * kLdsPitch];   // [k][v]
  __shared__ float sred[32][65];
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6, g = lane >> 4, r = lane & 15;
  const int kt = blockIdx.x;
  const int64_t ch = blockIdx.y;
  const int64_t vbase = ch * 32 * a.steps;
  const int64_t nxy = (int64_t)a.nx * a.ny, nvs = nxy * a.nz;
  // x-tile loader: voxel slot vs, k group q (8 consecutive k of this tile)
  const int vs = t >> 3, q = t & 7;
  const int kk0 = kt * 64 + 8 * q;
  const int tap = kk0 / a.cin, c0 = kk0 % a.cin;
  int dx, dy, dz;
  tap_offsets(tap, a.taps, dx, dy, dz);
  const bool kval = tap < a.taps;
  const int64_t off = dx + (int64_t)dy * a.nx + (int64_t)dz * nxy;
  // dy-tile loader
  const int gdy = a.cout >> 3;
  const bool dyl = t < 32 * gdy;
  const int dvs = dyl ? t / gdy : 0, dq = dyl ? t % gdy : 0;
  float bsum[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) bsum[i] = 0.f;
  f32x4_t acc[NBW];
#pragma unroll
  for (int nb = 0; nb < NBW; ++nb) acc[nb] = f32x4_t{0.f, 0.f, 0.f, 0.f};

  for (int s = 0; s < a.steps; ++s) {
    const int64_t v0 = vbase + 32 * s;
    {
      const int64_t v = v0 + vs;
      bool ok = kval && v < a.nvox;
      if (ok) {
        const int64_t loc = v % nvs;
        const int k = int(loc / nxy);
        const int64_t rem = loc - (int64_t)k * nxy;
        const int j = int(rem / a.nx), i = int(rem - (int64_t)j * a.nx);
        ok = unsigned(i + dx) < unsigned(a.nx) && unsigned(j + dy) < unsigned(a.ny) && unsigned(k + dz) < unsigned(a.nz);
      }
      u32x4_t u = {0u, 0u, 0u, 0u};
      if (ok) u = ld16(a.x + (v + off) * a.xpitch + c0);
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        sx[(8 * q + 2 * e) * kLdsPitch + vs] = uint16_t(u[e] & 0xffffu);
        sx[(8 * q + 2 * e + 1) * kLdsPitch + vs] = uint16_t(u[e] >> 16);
      }
    }
    if (dyl) {
      const int64_t v = v0 + dvs;
      u32x4_t u = {0u, 0u, 0u, 0u};
      if (v < a.nvox) u = ld16(a.dy + v * a.dypitch + 8 * dq);
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        sdy[(8 * dq + 2 * e) * kLdsPitch + dvs] = uint16_t(u[e] & 0xffffu);
        sdy[(8 * dq + 2 * e + 1) * kLdsPitch + dvs] = uint16_t(u[e] >> 16);
      }
      if (kt == 0) {
        float f[8];
        unpack8(u, f);
#pragma unroll
        for (int i = 0; i < 8; ++i) bsum[i] += f[i];
      }
    }
    __syncthreads();
    const bf16x8_t Bv = __builtin_bit_cast(bf16x8_t, ld16(&sx[(16 * wv + r) * kLdsPitch + 8 * g]));
#pragma unroll
    for (int nb = 0; nb < NBW; ++nb) {
      const bf16x8_t A = __builtin_bit_cast(bf16x8_t, ld16(&sdy[(16 * nb + r) * kLdsPitch + 8 * g]));
      acc[nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, Bv, acc[nb], 0, 0, 0);
    }
    __syncthreads();
  }
#pragma unroll
  for (int nb = 0; nb < NBW; ++nb)
#pragma unroll
    for (int i = 0; i < 4; ++i)
      a.part[(ch * a.cout + 16 * nb + 4 * g + i) * a.kp + kt * 64 + 16 * wv + r] = acc[nb][i];
  if (kt == 0) {  // bias gradient partial: sum over the chunk's voxels
    if (dyl) {
#pragma unroll
      for (int i = 0; i < 8; ++i) sred[dvs][8 * dq + i] = bsum[i];
    }
    __syncthreads();
    if (t < a.cout) {
      float sum = 0.f;
      for (int s = 0; s < 32; ++s) sum += sred[s][t];
      a.bpart[ch * a.cout + t] = sum;
    }
  }
}

__global__ void k_wgrad_reduce(const float* __restrict__ part, const float* __restrict__ bpart, int64_t nchunk,
                               int cout, int kp, float* __restrict__ gw, float* __restrict__ gb) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t per = (int64_t)cout * kp;
  if (idx < per) {
    float s = 0.f;
    for (int64_t c = 0; c < nchunk; ++c) s += part[c * per + idx];
    gw[idx] += s;
  } else if (idx < per + cout) {
    const int co = int(idx - per);
    float s = 0.f;
    for (int64_t c = 0; c < nchunk; ++c) s += bpart[c * cout + co];
    gb[co] += s;
  }
}

// master FP32 [cout_pad][kp] -> wt (bf16, same layout) and wd[ci][tap' * cout_pad + co]
__global__ void k_refresh(const float* __restrict__ w, int cout_pad, int kp, int cin, int taps, int kdp,
                          uint16_t* __restrict__ wt, uint16_t* __restrict__ wd) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nt = (int64_t)cout_pad * kp, nd = (int64_t)cin * kdp;
  if (idx < nt) wt[idx] = f2bf(w[idx]);
  if (idx < nd) {
    const int ci = int(idx / kdp), kk = int(idx % kdp);
    const int tp = kk / cout_pad, co = kk % cout_pad;
    float v = 0.f;
    if (tp < taps) v = w[(int64_t)co * kp + (int64_t)(taps - 1 - tp) * cin + ci];
    wd[idx] = f2bf(v);
  }
}

__global__ void k_adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m, float* __restrict__ v,
                       int64_t n, float lr, float b1, float b2, float eps, float c1, float c2, float wdecay) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  w[i] -= lr * ((mi / c1) / (sqrtf(vi / c2) + eps) + wdecay * w[i]);
}

struct XferK {
  const uint16_t* x;
  int xpitch;
  uint16_t* y;
  int ypitch, c;
  const uint16_t* mask;
  int mpitch;
  float s;
  int accum;
  int B, fx, fy, fz, cx, cy, cz;  // fine and coarse node dims per sample
};

__global__ void __launch_bounds__(256) k_net_restrict(XferK a) {
  const int G = a.c >> 3;
  const int64_t ncv = (int64_t)a.cx * a.cy * a.cz, nfv = (int64_t)a.fx * a.fy * a.fz;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)a.B * ncv * G) return;
  const int q = int(t % G);
  const int64_t v = t / G;
  const int b = int(v / ncv);
  const int64_t loc = v - (int64_t)b * ncv;
  const int K = int(loc / ((int64_t)a.cx * a.cy));
  const int64_t rem = loc - (int64_t)K * a.cx * a.cy;
  const int J = int(rem / a.cx), I = int(rem - (int64_t)J * a.cx);
  // all 27 loads are issued together: out-of-grid taps read a clamped
  // in-grid voxel with weight 0 (a loop that skipped them waited for each
  // load in turn: 12 us per launch whatever the grid size)
  u32x4_t raw[27];
  float wt[27];
#pragma unroll
  for (int dz = -1; dz <= 1; ++dz)
#pragma unroll
    for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
      for (int dx = -1; dx <= 1; ++dx) {
        const int t9 = (dz + 1) * 9 + (dy + 1) * 3 + (dx + 1);
        const int fk = 2 * K + dz, fj = 2 * J + dy, fi = 2 * I + dx;
        const bool in = fk >= 0 && fk < a.fz && fj >= 0 && fj < a.fy && fi >= 0 && fi < a.fx;
        const int ck = min(max(fk, 0), a.fz - 1), cj = min(max(fj, 0), a.fy - 1), ci = min(max(fi, 0), a.fx - 1);
        wt[t9] = in ? (dx ? 0.5f : 1.f) * (dy ? 0.5f : 1.f) * (dz ? 0.5f : 1.f) : 0.f;
        raw[t9] = ld16(a.x + ((int64_t)b * nfv + ci + (int64_t)a.fx * (cj + (int64_t)a.fy * ck)) * a.xpitch + 8 * q);
      }
  float acc[8];
#pragma unroll
  for (int e = 0; e < 8; ++e) acc[e] = 0.f;
#pragma unroll
  for (int t9 = 0; t9 < 27; ++t9) {
    float f[8];
    unpack8(raw[t9], f);
#pragma unroll
    for (int e = 0; e < 8; ++e) acc[e] += wt[t9] * f[e];
  }
  uint16_t* yp = a.y + v * a.ypitch + 8 * q;
  float o[8];
  if (a.accum) unpack8(ld16(yp), o);
#pragma unroll
  for (int e = 0; e < 8; ++e) o[e] = a.s * acc[e] + (a.accum ? o[e] : 0.f);
  if (a.mask) {
    float m[8];
    unpack8(ld16(a.mask + v * a.mpitch + 8 * q), m);
#pragma unroll
    for (int e = 0; e < 8; ++e)
      if (!(m[e] > 0.f)) o[e] = 0.f;
  }
  *reinterpret_cast<u32x4_t*>(yp) = pack8(o);
}

__global__ void __launch_bounds__(256) k_net_prolong(XferK a) {
  const int G = a.c >> 3;
  const int64_t ncv = (int64_t)a.cx * a.cy * a.cz, nfv = (int64_t)a.fx * a.fy * a.fz;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)a.B * nfv * G) return;
  const int q = int(t % G);
  const int64_t v = t / G;
  const int b = int(v / nfv);
  const int64_t loc = v - (int64_t)b * nfv;
  const int k = int(loc / ((int64_t)a.fx * a.fy));
  const int64_t rem = loc - (int64_t)k * a.fx * a.fy;
  const int j = int(rem / a.fx), i = int(rem - (int64_t)j * a.fx);
  int px[2], py[2], pz[2], nxp, nyp, nzp;
  float wx, wy, wz;
  // parents per axis: even -> one (weight 1), odd -> two (1/2 each)
  px[0] = i >> 1; px[1] = (i + 1) >> 1; nxp = (i & 1) ? 2 : 1; wx = (i & 1) ? 0.5f : 1.f;
  py[0] = j >> 1; py[1] = (j + 1) >> 1; nyp = (j & 1) ? 2 : 1; wy = (j & 1) ? 0.5f : 1.f;
  pz[0] = k >> 1; pz[1] = (k + 1) >> 1; nzp = (k & 1) ? 2 : 1; wz = (k & 1) ? 0.5f : 1.f;
  // the 8 parent loads are issued together; an axis with one parent reads it
  // twice, the second time with weight 0
  const float w = wx * wy * wz;
  u32x4_t raw[8];
  float wt[8];
#pragma unroll
  for (int c = 0; c < 2; ++c)
#pragma unroll
    for (int bb = 0; bb < 2; ++bb)
#pragma unroll
      for (int aa = 0; aa < 2; ++aa) {
        const int t8 = c * 4 + bb * 2 + aa;
        wt[t8] = (c < nzp && bb < nyp && aa < nxp) ? w : 0.f;
        raw[t8] = ld16(a.x + ((int64_t)b * ncv + px[aa] + (int64_t)a.cx * (py[bb] + (int64_t)a.cy * pz[c])) * a.xpitch +
                       8 * q);
      }
  float acc[8];
#pragma unroll
  for (int e = 0; e < 8; ++e) acc[e] = 0.f;
#pragma unroll
  for (int t8 = 0; t8 < 8; ++t8) {
    float f[8];
    unpack8(raw[t8], f);
#pragma unroll
    for (int e = 0; e < 8; ++e) acc[e] += wt[t8] * f[e];
  }
  uint16_t* yp = a.y + v * a.ypitch + 8 * q;
  float o[8];
  if (a.accum) unpack8(ld16(yp), o);
#pragma unroll
  for (int e = 0; e < 8; ++e) o[e] = a.s * acc[e] + (a.accum ? o[e] : 0.f);
  if (a.mask) {
    float m[8];
    unpack8(ld16(a.mask + v * a.mpitch + 8 * q), m);
#pragma unroll
    for (int e = 0; e < 8; ++e)
      if (!(m[e] > 0.f)) o[e] = 0.f;
  }
  *reinterpret_cast<u32x4_t*>(yp) = pack8(o);
}

// FP32 [v][c] -> bf16 [v][16] (zero padded)
__global__ void k_pad16(const float* __restrict__ d, int c, int64_t nvox, uint16_t* __restrict__ y) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nvox * 16) return;
  const int64_t v = t >> 4;
  const int ch = int(t & 15);
  y[t] = ch < c ? f2bf(d[v * c + ch]) : uint16_t(0);
}

constexpr int kLossBlocks = 64;

__global__ void __launch_bounds__(256) k_loss_partial(const float* __restrict__ o, const float* __restrict__ tg,
                                                      int64_t per, double* __restrict__ part) {
  const int b = blockIdx.y;
  double e = 0.0, s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < per; i += (int64_t)kLossBlocks * 256) {
    const double d = double(o[b * per + i]) - double(tg[b * per + i]);
    const double tv = tg[b * per + i];
    e += d * d;
    s += tv * tv;
  }
  __shared__ double se[256], ss[256];
  se[threadIdx.x] = e;
  ss[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      se[threadIdx.x] += se[threadIdx.x + w];
      ss[threadIdx.x] += ss[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[(b * kLossBlocks + blockIdx.x) * 2] = se[0];
    part[(b * kLossBlocks + blockIdx.x) * 2 + 1] = ss[0];
  }
}

__global__ void k_loss_final(const double* __restrict__ part, int B, double* __restrict__ inv_t, double* __restrict__ loss) {
  if (threadIdx.x != 0) return;
  double L = 0.0;
  for (int b = 0; b < B; ++b) {
    double e = 0.0, s = 0.0;
    for (int k = 0; k < kLossBlocks; ++k) {
      e += part[(b * kLossBlocks + k) * 2];
      s += part[(b * kLossBlocks + k) * 2 + 1];
    }
    const double it = s > 0.0 ? 1.0 / s : 0.0;
    inv_t[b] = it;
    L += e * it;
  }
  loss[0] = L / B;
}

__global__ void k_loss_grad(const float* __restrict__ o, const float* __restrict__ tg, int64_t per, int B,
                            const double* __restrict__ inv_t, float* __restrict__ dout) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= per * B) return;
  const int b = int(i / per);
  dout[i] = float(2.0 * (double(o[i]) - double(tg[i])) * inv_t[b] / B);
}

__global__ void k_sqnorm_partial(const float* __restrict__ g, int64_t n, double* __restrict__ acc) {
  __shared__ double sh[256];
  double s = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += 256) s += double(g[i]) * g[i];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) acc[0] += sh[0];
}

__global__ void k_scale(float* __restrict__ g, int64_t n, const double* __restrict__ s) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) g[i] = float(g[i] * s[0]);
}

inline unsigned blocks_for(int64_t n, int t = 256) { return unsigned((n + t - 1) / t); }

}  // namespace

// ---------------------------------------------------------------------------

void net_restrict(const View& x, const NetDims& f, const View& y, const NetDims& c, float s, const View* mask,
                  bool accum, hipStream_t st) {
  ANNEAL_CHECK(x.c == y.c && x.c % 8 == 0 && f.B == c.B, "net_restrict: channel or batch mismatch");
  XferK a{x.p, x.pitch, y.p, y.pitch, x.c, mask ? mask->p : nullptr, mask ? mask->pitch : 0, s, accum ? 1 : 0,
          f.B, f.nx, f.ny, f.nz, c.nx, c.ny, c.nz};
  const int64_t n = c.vox() * (x.c / 8);
  if (n == 0) return;
  hipLaunchKernelGGL(k_net_restrict, dim3(blocks_for(n)), dim3(256), 0, st, a);
  ANNEAL_HIP(hipGetLastError());
}

void net_prolong(const View& x, const NetDims& c, const View& y, const NetDims& f, float s, const View* mask,
                 bool accum, hipStream_t st) {
  ANNEAL_CHECK(x.c == y.c && x.c % 8 == 0 && f.B == c.B, "net_prolong: channel or batch mismatch");
  XferK a{x.p, x.pitch, y.p, y.pitch, x.c, mask ? mask->p : nullptr, mask ? mask->pitch : 0, s, accum ? 1 : 0,
          f.B, f.nx, f.ny, f.nz, c.nx, c.ny, c.nz};
  const int64_t n = f.vox() * (x.c / 8);
  if (n == 0) return;
  hipLaunchKernelGGL(k_net_prolong, dim3(blocks_for(n)), dim3(256), 0, st, a);
  ANNEAL_HIP(hipGetLastError());
}

void net_rel_mse(const float* o, const float* t, const NetDims& d, int c, float* dout, double* loss_dev, double* work,
                 hipStream_t st) {
  const int64_t per = (int64_t)d.nx * d.ny * d.nz * c;
  hipLaunchKernelGGL(k_loss_partial, dim3(kLossBlocks, d.B), dim3(256), 0, st, o, t, per, work);
  double* inv_t = work + 2 * kLossBlocks * d.B;
  hipLaunchKernelGGL(k_loss_final, dim3(1), dim3(64), 0, st, work, d.B, inv_t, loss_dev);
  if (dout) hipLaunchKernelGGL(k_loss_grad, dim3(blocks_for(per * d.B)), dim3(256), 0, st, o, t, per, d.B, inv_t, dout);
  ANNEAL_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------

void* UNet::alloc(size_t bytes) {
  void* p = nullptr;
  ANNEAL_HIP(hipMalloc(&p, std::max<size_t>(bytes, 256)));
  ANNEAL_HIP(hipMemsetAsync(p, 0, std::max<size_t>(bytes, 256), s_));
  owned_.push_back(p);
  return p;
}

static ConvLayer make_layer(int taps, int cin, int cout, bool relu) {
  ConvLayer L;
  L.taps = taps;
  L.cin = cin;
  L.cout = cout;
  L.cout_pad = (int)round_up(cout, 16);
  L.kp = (int)round_up((int64_t)taps * cin, 64);
  L.kdp = (int)round_up((int64_t)taps * L.cout_pad, 64);
  L.relu = relu;
  return L;
}

UNet::UNet(int cin, int cout, const std::vector<int>& widths, hipStream_t s)
    : s_(s), cin_(cin), cout_(cout), widths_(widths) {
  const int L = (int)widths.size();
  ANNEAL_CHECK(L >= 2 && L <= 6, "the U-Net needs 2..6 levels");
  ANNEAL_CHECK(cin >= 8 && cin % 8 == 0 && cin <= 128, "input channels must be a multiple of 8 (<= 128)");
  ANNEAL_CHECK(cout >= 1 && cout <= 16, "output channels must be 1..16");
  for (int w : widths) ANNEAL_CHECK(w % 16 == 0 && w >= 16 && w <= 64, "level widths must be 16, 32, 48 or 64");
  for (int l = 0; l < L; ++l) {
    enc_a_.push_back(make_layer(27, l == 0 ? cin : widths[l - 1], widths[l], true));
    enc_b_.push_back(make_layer(27, widths[l], widths[l], true));
  }
  for (int l = 0; l + 1 < L; ++l) {
    dec_a_.push_back(make_layer(27, widths[l + 1] + widths[l], widths[l], true));
    dec_b_.push_back(make_layer(27, widths[l], widths[l], true));
  }
  out_ = make_layer(1, widths[0], cout, false);
  auto init = [&](ConvLayer& Ly) {
    const size_t nw = (size_t)Ly.cout_pad * Ly.kp, nb = Ly.cout_pad;
    Ly.w = (float*)alloc(nw * 4);
    Ly.gw = (float*)alloc(nw * 4);
    Ly.mw = (float*)alloc(nw * 4);
    Ly.vw = (float*)alloc(nw * 4);
    Ly.b = (float*)alloc(nb * 4);
    Ly.gb = (float*)alloc(nb * 4);
    Ly.mb = (float*)alloc(nb * 4);
    Ly.vb = (float*)alloc(nb * 4);
    Ly.wt = (bf16*)alloc(nw * 2);
    Ly.wd = (bf16*)alloc((size_t)Ly.cin * Ly.kdp * 2);
  };
  for (auto& Ly : enc_a_) init(Ly);
  for (auto& Ly : enc_b_) init(Ly);
  for (auto& Ly : dec_a_) init(Ly);
  for (auto& Ly : dec_b_) init(Ly);
  init(out_);
  red_ = (double*)alloc(64 * sizeof(double));
  part_ = (double*)alloc((2 * kLossBlocks * 64 + 64) * sizeof(double));
  ANNEAL_HIP(hipStreamSynchronize(s_));
}

UNet::~UNet() {
  free_plan();
  if (kpart_) (void)hipFree(kpart_);
  for (void* p : owned_) (void)hipFree(p);
}

int64_t UNet::nparams() const {
  int64_t n = out_.nparams();
  for (auto* v : {&enc_a_, &enc_b_, &dec_a_, &dec_b_})
    for (auto& Ly : *v) n += Ly.nparams();
  return n;
}

template <typename F>
static void for_layers(std::vector<ConvLayer>& ea, std::vector<ConvLayer>& eb, std::vector<ConvLayer>& da,
                       std::vector<ConvLayer>& db, ConvLayer& o, F&& f) {
  for (size_t l = 0; l < ea.size(); ++l) {
    f(ea[l]);
    f(eb[l]);
  }
  for (size_t l = 0; l < da.size(); ++l) {
    f(da[l]);
    f(db[l]);
  }
  f(o);
}

void UNet::get_params(float* host) const {
  auto* self = const_cast<UNet*>(this);
  int64_t pos = 0;
  ANNEAL_HIP(hipStreamSynchronize(s_));
  for_layers(self->enc_a_, self->enc_b_, self->dec_a_, self->dec_b_, self->out_, [&](ConvLayer& Ly) {
    std::vector<float> w((size_t)Ly.cout_pad * Ly.kp), b(Ly.cout_pad);
    ANNEAL_HIP(hipMemcpy(w.data(), Ly.w, w.size() * 4, hipMemcpyDeviceToHost));
    ANNEAL_HIP(hipMemcpy(b.data(), Ly.b, b.size() * 4, hipMemcpyDeviceToHost));
    const int K = Ly.taps * Ly.cin;
    for (int co = 0; co < Ly.cout; ++co)
      for (int k = 0; k < K; ++k) host[pos++] = w[(size_t)co * Ly.kp + k];
    for (int co = 0; co < Ly.cout; ++co) host[pos++] = b[co];
  });
}

void UNet::get_grads(float* host) const {
  auto* self = const_cast<UNet*>(this);
  int64_t pos = 0;
  ANNEAL_HIP(hipStreamSynchronize(s_));
  for_layers(self->enc_a_, self->enc_b_, self->dec_a_, self->dec_b_, self->out_, [&](ConvLayer& Ly) {
    std::vector<float> w((size_t)Ly.cout_pad * Ly.kp), b(Ly.cout_pad);
    ANNEAL_HIP(hipMemcpy(w.data(), Ly.gw, w.size() * 4, hipMemcpyDeviceToHost));
    ANNEAL_HIP(hipMemcpy(b.data(), Ly.gb, b.size() * 4, hipMemcpyDeviceToHost));
    const int K = Ly.taps * Ly.cin;
    for (int co = 0; co < Ly.cout; ++co)
      for (int k = 0; k < K; ++k) host[pos++] = w[(size_t)co * Ly.kp + k];
    for (int co = 0; co < Ly.cout; ++co) host[pos++] = b[co];
  });
}

void UNet::set_params(const float* host) {
  int64_t pos = 0;
  for_layers(enc_a_, enc_b_, dec_a_, dec_b_, out_, [&](ConvLayer& Ly) {
    std::vector<float> w((size_t)Ly.cout_pad * Ly.kp, 0.f), b(Ly.cout_pad, 0.f);
    const int K = Ly.taps * Ly.cin;
    for (int co = 0; co < Ly.cout; ++co)
      for (int k = 0; k < K; ++k) w[(size_t)co * Ly.kp + k] = host[pos++];
    for (int co = 0; co < Ly.cout; ++co) b[co] = host[pos++];
    ANNEAL_HIP(hipMemcpy(Ly.w, w.data(), w.size() * 4, hipMemcpyHostToDevice));
    ANNEAL_HIP(hipMemcpy(Ly.b, b.data(), b.size() * 4, hipMemcpyHostToDevice));
    const size_t nw = w.size() * 4, nb = b.size() * 4;
    ANNEAL_HIP(hipMemsetAsync(Ly.mw, 0, nw, s_));
    ANNEAL_HIP(hipMemsetAsync(Ly.vw, 0, nw, s_));
    ANNEAL_HIP(hipMemsetAsync(Ly.mb, 0, nb, s_));
    ANNEAL_HIP(hipMemsetAsync(Ly.vb, 0, nb, s_));
    refresh_bf16(Ly);
  });
  zero_grad();
  ANNEAL_HIP(hipStreamSynchronize(s_));
}

void UNet::refresh_bf16(ConvLayer& Ly) {
  const int64_t n = std::max<int64_t>((int64_t)Ly.cout_pad * Ly.kp, (int64_t)Ly.cin * Ly.kdp);
  hipLaunchKernelGGL(k_refresh, dim3(blocks_for(n)), dim3(256), 0, s_, Ly.w, Ly.cout_pad, Ly.kp, Ly.cin, Ly.taps,
                     Ly.kdp, Ly.wt, Ly.wd);
  ANNEAL_HIP(hipGetLastError());
}

void UNet::zero_grad() {
  for_layers(enc_a_, enc_b_, dec_a_, dec_b_, out_, [&](ConvLayer& Ly) {
    ANNEAL_HIP(hipMemsetAsync(Ly.gw, 0, (size_t)Ly.cout_pad * Ly.kp * 4, s_));
    ANNEAL_HIP(hipMemsetAsync(Ly.gb, 0, (size_t)Ly.cout_pad * 4, s_));
  });
}

void UNet::adam(float lr, float b1, float b2, float eps, int step, float wdecay) {
  ANNEAL_CHECK(step >= 1, "adam step counts from 1");
  const float c1 = 1.f - std::pow(b1, (float)step), c2 = 1.f - std::pow(b2, (float)step);
  for_layers(enc_a_, enc_b_, dec_a_, dec_b_, out_, [&](ConvLayer& Ly) {
    const int64_t nw = (int64_t)Ly.cout_pad * Ly.kp, nb = Ly.cout_pad;
    hipLaunchKernelGGL(k_adam, dim3(blocks_for(nw)), dim3(256), 0, s_, Ly.w, Ly.gw, Ly.mw, Ly.vw, nw, lr, b1, b2, eps,
                       c1, c2, wdecay);
    hipLaunchKernelGGL(k_adam, dim3(blocks_for(nb)), dim3(256), 0, s_, Ly.b, Ly.gb, Ly.mb, Ly.vb, nb, lr, b1, b2, eps,
                       c1, c2, 0.f);
    refresh_bf16(Ly);
  });
  ANNEAL_HIP(hipGetLastError());
}

void UNet::grad_sqnorm(double* out_dev) {
  ANNEAL_HIP(hipMemsetAsync(out_dev, 0, sizeof(double), s_));
  for_layers(enc_a_, enc_b_, dec_a_, dec_b_, out_, [&](ConvLayer& Ly) {
    hipLaunchKernelGGL(k_sqnorm_partial, dim3(1), dim3(256), 0, s_, Ly.gw, (int64_t)Ly.cout_pad * Ly.kp, out_dev);
    hipLaunchKernelGGL(k_sqnorm_partial, dim3(1), dim3(256), 0, s_, Ly.gb, (int64_t)Ly.cout_pad, out_dev);
  });
  ANNEAL_HIP(hipGetLastError());
}

void UNet::scale_grad(const double* s) {
  for_layers(enc_a_, enc_b_, dec_a_, dec_b_, out_, [&](ConvLayer& Ly) {
    const int64_t nw = (int64_t)Ly.cout_pad * Ly.kp;
    hipLaunchKernelGGL(k_scale, dim3(blocks_for(nw)), dim3(256), 0, s_, Ly.gw, nw, s);
    hipLaunchKernelGGL(k_scale, dim3(1), dim3(256), 0, s_, Ly.gb, (int64_t)Ly.cout_pad, s);
  });
  ANNEAL_HIP(hipGetLastError());
}

// ---- plan (activation and gradient buffers for one input shape) ------------

void UNet::free_plan() {
  if (!have_plan_) return;
  ANNEAL_HIP(hipStreamSynchronize(s_));
  for (bf16* p : plan_.owned) (void)hipFree(p);
  if (plan_.wpart) (void)hipFree(plan_.wpart);
  plan_ = Plan{};
  have_plan_ = false;
  fwd_valid_ = false;
}

void UNet::make_plan(const NetDims& d0) {
  free_plan();
  Plan& P = plan_;
  const int L = levels();
  P.d[0] = d0;
  for (int l = 1; l < L; ++l)
    P.d[l] = NetDims{d0.B, coarse_n(P.d[l - 1].nx), coarse_n(P.d[l - 1].ny), coarse_n(P.d[l - 1].nz)};
  auto buf = [&](int64_t vox, int c) -> View {
    void* p = nullptr;
    const size_t bytes = std::max<size_t>((size_t)vox * c * 2, 256);
    ANNEAL_HIP(hipMalloc(&p, bytes));
    ANNEAL_HIP(hipMemsetAsync(p, 0, bytes, s_));
    P.owned.push_back((bf16*)p);
    return View{(bf16*)p, c, c};
  };
  for (int l = 0; l < L; ++l) {
    const int64_t V = P.d[l].vox();
    const int w = widths_[l];
    if (l > 0) {
      P.x[l] = buf(V, widths_[l - 1]);
      P.gx[l] = buf(V, widths_[l - 1]);
    }
    P.a[l] = buf(V, w);
    P.ga[l] = buf(V, w);
    if (l + 1 < L) {
      const int up = widths_[l + 1];
      P.cat[l] = buf(V, up + w);
      P.gcat[l] = buf(V, up + w);
      P.e[l] = buf(V, w);
      P.ge[l] = buf(V, w);
      P.dec[l] = buf(V, w);
      P.gdec[l] = buf(V, w);
    } else {
      P.bott = buf(V, w);
      P.gbott = buf(V, w);
    }
  }
  {
    void* p = nullptr;
    const size_t bytes = std::max<size_t>((size_t)d0.vox() * 16 * 2, 256);
    ANNEAL_HIP(hipMalloc(&p, bytes));
    P.owned.push_back((bf16*)p);
    P.dout16 = (bf16*)p;
  }
  have_plan_ = true;
  fwd_valid_ = false;
}

static void pick_nb(int n16, int& NB, int& MB) {
  NB = (n16 % 4 == 0) ? 4 : (n16 % 3 == 0) ? 3 : (n16 % 2 == 0) ? 2 : 1;
  MB = NB >= 3 ? 2 : 4;
}

template <int NB, int MB>
static void launch_conv(const ConvK& a, int n16, hipStream_t s) {
  const int64_t per_block = 64 * MB;
  dim3 grid(unsigned((a.nvox + per_block - 1) / per_block), unsigned(n16 / NB), unsigned(a.ksplit));
  hipLaunchKernelGGL((k_conv<NB, MB>), grid, dim3(256), 0, s, a);
}

constexpr int kBrickMB = 4;  // z layers per brick

template <int S>
static void launch_brick(const ConvK& a, int n16, int64_t bricks, hipStream_t s) {
  // all output channels of a brick in one workgroup when they fit (up to 64), else 64 per workgroup;
  // on grids with few bricks the output channels are split further until there are at least two
  // workgroups per CU (the 49x25x25 level has 196 bricks: one workgroup per brick left a quarter of
  // the CUs idle and ran each brick's whole output width on one CU)
  int NB = n16 >= 4 ? 4 : n16;
  while (NB > 1 && bricks * (n16 / NB) < 512) NB /= 2;
  dim3 grid(unsigned(bricks), unsigned(n16 / NB));
  if (NB == 4) hipLaunchKernelGGL((k_conv_brick<4, S, kBrickMB>), grid, dim3(256), 0, s, a);
  else if (NB == 2) hipLaunchKernelGGL((k_conv_brick<2, S, kBrickMB>), grid, dim3(256), 0, s, a);
  else hipLaunchKernelGGL((k_conv_brick<1, S, kBrickMB>), grid, dim3(256), 0, s, a);
}

// 3x3x3 convolutions on grids of at least kBrickMin bricks run k_conv_brick
// (input staged in LDS); the others, and the 1x1x1 output convolution, run
// k_conv, whose K loop is split over blockIdx.z on grids smaller than the
// machine (at least 4 chunks each) so enough waves overlap their load latency.
constexpr int64_t kBrickMin = 128;

void UNet::conv_dispatch(ConvK a, int n16) {
  if (a.taps == 27 && (n16 == 1 || n16 == 2 || n16 % 4 == 0) && a.cin % 8 == 0 && a.kp >= 27 * a.cin) {
    const int64_t nvs = (int64_t)a.nx * a.ny * a.nz;
    const int64_t bricks = (a.nvox / nvs) * ((a.nx + kBX - 1) / kBX) * ((a.ny + kBY - 1) / kBY) *
                           ((a.nz + kBrickMB - 1) / kBrickMB);
    if (bricks >= kBrickMin) {
      if (a.cin % 32 == 0) launch_brick<32>(a, n16, bricks, s_);
      else if (a.cin % 16 == 0) launch_brick<16>(a, n16, bricks, s_);
      else launch_brick<8>(a, n16, bricks, s_);
      ANNEAL_HIP(hipGetLastError());
      return;
    }
  }
  int NB, MB;
  pick_nb(n16, NB, MB);
  const int64_t blocks = ((a.nvox + 64 * MB - 1) / (64 * MB)) * (n16 / NB);
  a.ksplit = 1;
  a.kper = a.nchunk;
  a.nout = n16 * 16;
  if (blocks < 512 && a.nchunk >= 8) {
    int ks = (int)std::min<int64_t>(a.nchunk / 4, (1024 + blocks - 1) / blocks);
    if (ks > 1) {
      a.kper = (a.nchunk + ks - 1) / ks;
      a.ksplit = (a.nchunk + a.kper - 1) / a.kper;
    }
  }
  if (a.ksplit > 1) {
    const size_t need = (size_t)a.ksplit * a.nvox * a.nout;
    if (need > kpart_elems_) {
      if (kpart_) ANNEAL_HIP(hipFree(kpart_));
      kpart_ = nullptr;
      ANNEAL_HIP(hipMalloc(&kpart_, need * sizeof(float)));
      kpart_elems_ = need;
    }
    a.part = kpart_;
  }
  if (NB == 4) launch_conv<4, 2>(a, n16, s_);
  else if (NB == 3) launch_conv<3, 2>(a, n16, s_);
  else if (NB == 2) launch_conv<2, 4>(a, n16, s_);
  else launch_conv<1, 4>(a, n16, s_);
  if (a.ksplit > 1) {
    const int64_t n = a.nvox * a.nout;
    hipLaunchKernelGGL(k_conv_finish, dim3(blocks_for(n)), dim3(256), 0, s_, a);
  }
  ANNEAL_HIP(hipGetLastError());
}

void UNet::conv_fwd(ConvLayer& Ly, const View& x, const NetDims& d, const View& y, float* yf, int cf) {
  ANNEAL_CHECK(x.c == Ly.cin, "conv input channels do not match the layer");
  ConvK a{};
  a.x = x.p;
  a.xpitch = x.pitch;
  a.cin = Ly.cin;
  a.taps = Ly.taps;
  a.w = Ly.wt;
  a.kp = Ly.kp;
  a.nchunk = int((Ly.taps * Ly.cin + 31) / 32);
  a.bias = Ly.b;
  a.y = y.p;
  a.ypitch = y.pitch;
  a.yf = yf;
  a.cf = cf;
  a.relu = Ly.relu ? 1 : 0;
  a.nx = d.nx;
  a.ny = d.ny;
  a.nz = d.nz;
  a.nvox = d.vox();
  conv_dispatch(a, Ly.cout_pad / 16);
}

void UNet::conv_dgrad(ConvLayer& Ly, const View& dy, const NetDims& d, const View& dx, const View* mask, bool accum) {
  ANNEAL_CHECK(dy.c == Ly.cout_pad && dx.c == Ly.cin && Ly.cin % 16 == 0, "dgrad view mismatch");
  ConvK a{};
  a.x = dy.p;
  a.xpitch = dy.pitch;
  a.cin = Ly.cout_pad;
  a.taps = Ly.taps;
  a.w = Ly.wd;
  a.kp = Ly.kdp;
  a.nchunk = int((Ly.taps * Ly.cout_pad + 31) / 32);
  a.y = dx.p;
  a.ypitch = dx.pitch;
  a.mask = mask ? mask->p : nullptr;
  a.mpitch = mask ? mask->pitch : 0;
  a.accum = accum ? 1 : 0;
  a.nx = d.nx;
  a.ny = d.ny;
  a.nz = d.nz;
  a.nvox = d.vox();
  conv_dispatch(a, Ly.cin / 16);
}

void UNet::conv_wgrad(ConvLayer& Ly, const View& x, const View& dy, const NetDims& d) {
  ANNEAL_CHECK(x.c == Ly.cin && dy.c == Ly.cout_pad && Ly.cout_pad <= 64, "wgrad view mismatch");
  const int steps = 64;  // 2048 voxels per chunk
  const int64_t nvox = d.vox();
  const int64_t nchunk = (nvox + 32 * steps - 1) / (32 * steps);
  const size_t need = (size_t)nchunk * Ly.cout_pad * (Ly.kp + 1);
  if (need > plan_.wpart_elems) {
    if (plan_.wpart) ANNEAL_HIP(hipFree(plan_.wpart));
    ANNEAL_HIP(hipMalloc(&plan_.wpart, need * sizeof(float)));
    plan_.wpart_elems = need;
  }
  WgradK a{};
  a.x = x.p;
  a.xpitch = x.pitch;
  a.cin = Ly.cin;
  a.taps = Ly.taps;
  a.dy = dy.p;
  a.dypitch = dy.pitch;
  a.cout = Ly.cout_pad;
  a.part = plan_.wpart;
  a.bpart = plan_.wpart + (size_t)nchunk * Ly.cout_pad * Ly.kp;
  a.kp = Ly.kp;
  a.nx = d.nx;
  a.ny = d.ny;
  a.nz = d.nz;
  a.nvox = nvox;
  a.steps = steps;
  dim3 grid(unsigned(Ly.kp / 64), unsigned(nchunk));
  const int nbw = Ly.cout_pad / 16;
  if (nbw == 4) hipLaunchKernelGGL(k_wgrad<4>, grid, dim3(256), 0, s_, a);
  else if (nbw == 3) hipLaunchKernelGGL(k_wgrad<3>, grid, dim3(256), 0, s_, a);
  else if (nbw == 2) hipLaunchKernelGGL(k_wgrad<2>, grid, dim3(256), 0, s_, a);
  else hipLaunchKernelGGL(k_wgrad<1>, grid, dim3(256), 0, s_, a);
  const int64_t nred = (int64_t)Ly.cout_pad * Ly.kp + Ly.cout_pad;
  hipLaunchKernelGGL(k_wgrad_reduce, dim3(blocks_for(nred)), dim3(256), 0, s_, a.part, a.bpart, nchunk, Ly.cout_pad,
                     Ly.kp, Ly.gw, Ly.gb);
  ANNEAL_HIP(hipGetLastError());
}

void UNet::forward(const bf16* in, const NetDims& d0, float* out) {
  ANNEAL_CHECK(d0.B >= 1 && d0.nx >= 1 && d0.ny >= 1 && d0.nz >= 1, "empty input");
  if (!have_plan_ || !(plan_.d[0] == d0)) make_plan(d0);
  Plan& P = plan_;
  const int L = levels();
  P.in = View{const_cast<bf16*>(in), cin_, cin_};
  P.out = out;
  // encoder
  for (int l = 0; l < L; ++l) {
    const View xin = l == 0 ? P.in : P.x[l];
    conv_fwd(enc_a_[l], xin, P.d[l], P.a[l], nullptr, 0);
    const View s = l + 1 < L ? P.cat[l].sub(widths_[l + 1], widths_[l]) : P.bott;
    conv_fwd(enc_b_[l], P.a[l], P.d[l], s, nullptr, 0);
    if (l + 1 < L) net_restrict(s, P.d[l], P.x[l + 1], P.d[l + 1], 0.125f, nullptr, false, s_);
  }
  // decoder
  for (int l = L - 2; l >= 0; --l) {
    const View below = l + 1 == L - 1 ? P.bott : P.dec[l + 1];
    net_prolong(below, P.d[l + 1], P.cat[l].sub(0, widths_[l + 1]), P.d[l], 1.f, nullptr, false, s_);
    conv_fwd(dec_a_[l], P.cat[l], P.d[l], P.e[l], nullptr, 0);
    conv_fwd(dec_b_[l], P.e[l], P.d[l], P.dec[l], nullptr, 0);
  }
  conv_fwd(out_, P.dec[0], P.d[0], View{}, out, cout_);
  fwd_valid_ = true;
}

void UNet::backward(const float* dout) {
  ANNEAL_CHECK(fwd_valid_, "backward needs a forward pass of the same shape first");
  Plan& P = plan_;
  const int L = levels();
  const NetDims& d0 = P.d[0];
  hipLaunchKernelGGL(k_pad16, dim3(blocks_for(d0.vox() * 16)), dim3(256), 0, s_, dout, cout_, d0.vox(), P.dout16);
  const View g_out{P.dout16, 16, 16};
  conv_wgrad(out_, P.dec[0], g_out, d0);
  conv_dgrad(out_, g_out, d0, P.gdec[0], &P.dec[0], false);
  // decoder, top level first
  for (int l = 0; l + 1 < L; ++l) {
    conv_wgrad(dec_b_[l], P.e[l], P.gdec[l], P.d[l]);
    conv_dgrad(dec_b_[l], P.gdec[l], P.d[l], P.ge[l], &P.e[l], false);
    conv_wgrad(dec_a_[l], P.cat[l], P.ge[l], P.d[l]);
    conv_dgrad(dec_a_[l], P.ge[l], P.d[l], P.gcat[l], nullptr, false);
    const bool bottom = l + 1 == L - 1;
    const View below = bottom ? P.bott : P.dec[l + 1];
    const View gbelow = bottom ? P.gbott : P.gdec[l + 1];
    net_restrict(P.gcat[l].sub(0, widths_[l + 1]), P.d[l], gbelow, P.d[l + 1], 1.f, &below, false, s_);
  }
  // encoder, bottom level first
  for (int l = L - 1; l >= 0; --l) {
    const View s = l + 1 < L ? P.cat[l].sub(widths_[l + 1], widths_[l]) : P.bott;
    const View gs = l + 1 < L ? P.gcat[l].sub(widths_[l + 1], widths_[l]) : P.gbott;
    conv_wgrad(enc_b_[l], P.a[l], gs, P.d[l]);
    conv_dgrad(enc_b_[l], gs, P.d[l], P.ga[l], &P.a[l], false);
    const View xin = l == 0 ? P.in : P.x[l];
    conv_wgrad(enc_a_[l], xin, P.ga[l], P.d[l]);
    if (l > 0) {
      conv_dgrad(enc_a_[l], P.ga[l], P.d[l], P.gx[l], nullptr, false);
      const View sp = P.cat[l - 1].sub(widths_[l], widths_[l - 1]);
      const View gsp = P.gcat[l - 1].sub(widths_[l], widths_[l - 1]);
      net_prolong(P.gx[l], P.d[l], gsp, P.d[l - 1], 0.125f, &sp, true, s_);
    }
  }
}

}  // namespace anneal
