// Matrix-free hex8 operator on the voxel grid and the geometric-multigrid
// transfer kernels (gfx950).
//
// The operator replaces the assembled CSR of anneal/fem.py:156-201: K u =
// sum_e Ee_e * ke * u_e, with ke the 2x2x2-Gauss hex8 stiffness (fem.py:71-84).
//
// Apply kernel design (z-marching, owner-computes, no atomics):
//   * a workgroup owns a TX x TY tile of element columns (one thread each) and
//     walks a chunk of ZC node planes in +z;
//   * each node plane of u is staged in LDS once; a thread forms the 2-D
//     Walsh (sum/difference) transform of its column's 4 corner nodes, so an
//     element's 8-corner transform is one add/sub per mode of two consecutive
//     planes (register carried);
//   * in the Walsh basis ke has 45 non-zeros in tiny blocks (about 33 FMAs per
//     element instead of 576);
//   * the face-domain outputs of the element below and above a node plane are
//     summed in registers, transformed back, written to LDS, and each node
//     gathers its 4 element columns in a fixed order: bit-reproducible.
// The tile has one halo element column on the low x and y sides, so it
// produces (TX-1) x (TY-1) node columns.
#include "stencil.hpp"

namespace anneal {

template <typename T>
__device__ __forceinline__ KeCoef<T> scaled(const KeCoef<T>& k, T e) {
  return KeCoef<T>{k.nAmB * e, k.nB * e, k.g * e, k.pc * e, k.pd * e, k.tAmB * e, k.tB * e, k.w * e};
}

// Raw buffer access.  A resource covers one whole array (the range is its
// byte length, below 2 GiB) and each access adds a uniform byte offset (the
// node plane or element layer, in a scalar register) to the lane's offset.
// Lanes outside the tile's valid region use the offset kOob = 2^31, so their
// loads return 0 and their stores are dropped by the hardware range check,
// with no per-lane selects or branches (kOob plus any plane offset below 2^31
// stays out of range, whether or not the check includes the scalar offset).
constexpr uint32_t kOob = 0x80000000u;
__device__ __forceinline__ __amdgpu_buffer_rsrc_t rsrc(const void* p, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), 0, bytes, 0x00020000);
}
template <typename T>
__device__ __forceinline__ T bload(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
  if constexpr (sizeof(T) == 4) {
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
  } else {
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
  }
}
template <typename T>
__device__ __forceinline__ void bstore(T v, __amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
  if constexpr (sizeof(T) == 4) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, voff, soff, 0);
  } else {
    using V2 = __attribute__((ext_vector_type(2))) uint32_t;
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(V2, v), r, voff, soff, 0);
  }
}


// FP32 variants are held to 5 waves per SIMD (at most 96 VGPRs with gfx950's
// 8-register granule): at 97-104 VGPRs the smoothing modes ran at 4 waves, so
// the 1274 workgroups of the 192x96x96 fine level no longer fit in one round
// on 256 CUs (25-29 us instead of 20-23 us per launch).
template <typename T, int TX, int TY, int ZC, int MODE>
__global__ __launch_bounds__(TX* TY, sizeof(T) == 4 ? 5 : 1) void k_apply(ApplyArgs<T> a) {
  constexpr int LX = TX + 1, LY = TY + 1, NT = TX * TY, NP = LX * LY;
  constexpr int NLD = (NP + NT - 1) / NT;  // node loads per thread per plane
  __shared__ T su[3][3][NP];               // three planes: p is written while p-1 is still read
  __shared__ T sp[12][NT];
  __shared__ double sred[NT / 64];

  const int tid = threadIdx.x;
  const int tx = tid % TX, ty = tid / TX;
  const int nb = a.ntx * a.nty * a.nzc;
  // Workgroups reach the 8 XCDs round-robin in launch order; xcd_logical gives
  // each XCD a contiguous run of (rhs, tile) ids, and tiles are numbered x
  // fastest, then y, then z, so an XCD holds whole rows of the tile grid:
  // their shared halo columns and rows, and the cache lines that two x
  // neighbours both write in part, meet in one L2 (z fastest, the former
  // order, was 3 % slower over the SIMP step; no remapping, 20 %).
  const int gid = xcd_logical(blockIdx.x + nb * blockIdx.y, nb * gridDim.y);
  const int r = gid / nb;
  const int lid = gid - r * nb;
  const int bx = lid % a.ntx;
  const int byz = lid / a.ntx;
  const int by = byz % a.nty, bz = byz / a.nty;
  if (rhs_skipped(a.skip, r)) return;
  const int X0 = bx * (TX - 1), Y0 = by * (TY - 1), Z0 = bz * ZC;
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const int sx = nx + 1;
  const int64_t sxy = (int64_t)(nx + 1) * (ny + 1);
  const int64_t ldn = a.ldn;
  const T* __restrict__ u = a.u + (int64_t)3 * r * ldn;
  T* y = a.y + (int64_t)3 * r * ldn;  // not restrict: kCheb may write over prev
  const T* __restrict__ bb = (MODE != kApply) ? a.b + (int64_t)3 * r * ldn : nullptr;
  const T* pr = (MODE == kCheb && a.prev) ? a.prev + (int64_t)3 * r * ldn : nullptr;
  T* po = MODE == kApplyDir ? a.pout + (int64_t)3 * r * ldn : nullptr;  // the new direction
  // kApplyDir: the staged u is the direction p = b + beta u_old (as the PCG's separate direction update)
  const T beta = MODE == kApplyDir ? T(a.skip[r * kSt + kBeta]) : T(0);
  // operator of this right-hand side (one problem of a group, or the only one)
  const T* opEe = a.Ee;
  const uint8_t* opmask = a.mask;
  const float* opdinv = a.dinv;
  if (a.ops) {
    const OpEntry& o = a.ops[r / a.rpo];
    if constexpr (sizeof(T) == 8) opEe = (const T*)o.Ee64;
    else opEe = (const T*)o.Ee;
    opmask = o.mask;
    opdinv = o.dinv;
  }

  // Addressing: every access goes through the buffer resource of its array
  // (uniform, scalar registers), the plane's byte offset (scalar) and a
  // per-thread byte offset that is fixed along z; out-of-range lanes carry kOob.
  uint32_t lvo[NLD];  // byte offsets of this thread's load slots in a node plane
#pragma unroll
  for (int k = 0; k < NLD; ++k) {
    const int idx = tid + k * NT;
    const int lx = idx % LX, ly = idx / LX;
    const int gx = X0 - 1 + lx, gy = Y0 - 1 + ly;
    const bool ok = idx < NP && gx >= 0 && gx <= nx && gy >= 0 && gy <= ny;
    lvo[k] = ok ? uint32_t(gx + sx * gy) * uint32_t(sizeof(T)) : kOob;
  }
  const int ei = X0 - 1 + tx, ej = Y0 - 1 + ty;
  const bool ecol = (ei >= 0) && (ei < nx) && (ej >= 0) && (ej < ny);
  const uint32_t evo = ecol ? uint32_t(ei + nx * ej) * uint32_t(sizeof(T)) : kOob;
  const int64_t exy = (int64_t)nx * ny;
  const int ni = X0 + tx, nj = Y0 + ty;
  const bool ncol = (tx < TX - 1) && (ty < TY - 1) && (ni <= nx) && (nj <= ny);
  const uint32_t nvo = ncol ? uint32_t(ni + sx * nj) : kOob;   // element index (scaled per array below)
  const uint32_t pbytes = uint32_t(sxy) * uint32_t(sizeof(T));

  // Plane addressing: a plane's byte offset in its array (32 bits: the
  // launcher keeps the 3 ldn entries of one right-hand side below 2 GiB).
  // Planes outside [0, nz] are clamped: their values only meet element layers
  // outside the grid, whose stiffness is 0, and no output is stored for them.
  auto plane_off = [&](int p) -> uint32_t { return uint32_t(min(max(p, 0), nz)) * pbytes; };
  // components of a node array lie ldn entries apart
  const uint32_t cstep = uint32_t(ldn) * uint32_t(sizeof(T));
  const uint32_t vbytes = 3u * cstep;                                       // one rhs of a node array
  const uint32_t nnode = uint32_t(sxy) * uint32_t(nz + 1);
  const __amdgpu_buffer_rsrc_t ru = rsrc(u, vbytes), ry = rsrc(y, vbytes);
  const __amdgpu_buffer_rsrc_t rb = rsrc(bb, MODE != kApply ? vbytes : 0u);
  const __amdgpu_buffer_rsrc_t rp = rsrc(pr ? pr : u, pr ? vbytes : 0u);  // empty: the loads return 0
  const __amdgpu_buffer_rsrc_t rpo = rsrc(po, po ? vbytes : 0u);
  const __amdgpu_buffer_rsrc_t rm = rsrc(opmask, nnode);
  const __amdgpu_buffer_rsrc_t rd = rsrc(opdinv, (MODE == kJacobi || MODE == kCheb || MODE == kCheb0) ? nnode * 4u : 0u);

  // Node planes are prefetched two steps ahead (two register slots), so the
  // load latency of a plane is covered by two steps of compute.
  // (the fused first smoothing steps stage mask, D^-1 and b per node, and the
  // direction mode two arrays, so they prefetch one plane ahead to stay within
  // 96 VGPRs)
  constexpr int kDepth = (MODE == kCheb0 || MODE == kApplyDir) ? 1 : 2;
  constexpr int kDir = MODE == kApplyDir ? 1 : 0;  // a second staged array (b) for the direction update
  T pre[kDepth][NLD][3];
  T preb[kDir ? kDepth : 1][NLD][3];
  auto fetch = [&](int p, int slot) {
    const uint32_t o = plane_off(p);
    if (MODE == kCheb0) {
      // first smoothing step from a zero guess: u = M omega D^-1 b
      const uint32_t on = o / uint32_t(sizeof(T));  // the plane's first node
#pragma unroll
      for (int k = 0; k < NLD; ++k) {
        const uint32_t node = lvo[k] == kOob ? kOob : lvo[k] / uint32_t(sizeof(T));
        const uint8_t m = __builtin_amdgcn_raw_buffer_load_b8(rm, node, on, 0);
        const T d = a.omega * T(bload<float>(rd, node == kOob ? kOob : node * 4u, on * 4u));
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const T bv = bload<T>(rb, lvo[k], o + c * cstep);
          pre[slot][k][c] = ((m >> c) & 1) ? d * bv : T(0);
        }
      }
    } else {
#pragma unroll
      for (int k = 0; k < NLD; ++k) {
        pre[slot][k][0] = bload<T>(ru, lvo[k], o);
        pre[slot][k][1] = bload<T>(ru, lvo[k], o + cstep);
        pre[slot][k][2] = bload<T>(ru, lvo[k], o + 2 * cstep);
        if (kDir) {
#pragma unroll
          for (int c = 0; c < 3; ++c) preb[kDir ? slot : 0][k][c] = bload<T>(rb, lvo[k], o + c * cstep);
        }
      }
    }
  };
  const uint32_t ebytes = uint32_t(nx) * uint32_t(ny) * uint32_t(sizeof(T));  // one element layer
  const __amdgpu_buffer_rsrc_t re = rsrc(opEe, ebytes * uint32_t(nz));
  auto fetch_e = [&](int L) -> T {
    const bool lin = L >= 0 && L < nz;  // layers outside the grid read as 0
    return bload<T>(re, lin ? evo : kOob, uint32_t(min(max(L, 0), nz - 1)) * ebytes);
  };

  // Operands of the output node in plane L (mask, b, D^-1, previous iterate),
  // also fetched one step before they are used (planes outside the chunk are
  // clamped and their operands unused).
  struct Epi {
    uint8_t mk;
    float dv;
    T bv[3], pv[3];
  };
  const uint32_t ovo = nvo == kOob ? kOob : nvo * uint32_t(sizeof(T));
  const uint32_t dvo = nvo == kOob ? kOob : nvo * 4u;
  auto fetch_epi = [&](int L, Epi& E) {
    const uint32_t o = plane_off(L);
    const uint32_t on = o / uint32_t(sizeof(T));
    E.mk = __builtin_amdgcn_raw_buffer_load_b8(rm, nvo, on, 0);
    E.dv = 0.f;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      E.bv[c] = T(0);
      E.pv[c] = T(0);
    }
    if (MODE != kApply && MODE != kApplyDir) {
#pragma unroll
      for (int c = 0; c < 3; ++c) E.bv[c] = bload<T>(rb, ovo, o + c * cstep);
    }
    if (MODE == kJacobi || MODE == kCheb || MODE == kCheb0) E.dv = bload<float>(rd, dvo, on * 4u);
    if (MODE == kCheb) {
      // without a previous-iterate array the resource is empty and the loads
      // return 0 (no branch: merging loaded and computed values would wait
      // for the loads right here)
#pragma unroll
      for (int c = 0; c < 3; ++c) E.pv[c] = bload<T>(rp, ovo, o + c * cstep);
    }
  };

  T Fprev[3][4], gup[3][4];
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      Fprev[c][m] = T(0);
      gup[c][m] = T(0);
    }
  double dot = 0.0;
  fetch(Z0 - 1, 0);
  if (kDepth == 2) fetch(Z0, 1 % kDepth);
  T eq[2] = {fetch_e(Z0 - 1), fetch_e(Z0)};  // layers used at steps 1 and 2
  // Light modes (K u, residual) fetch the epilogue operands one step ahead;
  // the smoothing modes read four more arrays there, and a second operand set
  // would push them past 96 VGPRs (4 instead of 5 waves per SIMD), so they
  // fetch at the top of the step that uses them (the step's compute covers
  // the latency).
  constexpr bool kEpiAhead = (MODE == kApply || MODE == kResid || MODE == kApplyDir);
  Epi ep[kEpiAhead ? 2 : 1];
  if (kEpiAhead) fetch_epi(Z0, ep[0]);  // plane finished at step 2

#pragma unroll
  for (int s = 0; s <= ZC + 1; ++s) {
    const int p = Z0 - 1 + s;  // node plane staged in this step
    T* buf = &su[s % 3][0][0];
#pragma unroll
    for (int k = 0; k < NLD; ++k) {
      const int idx = tid + k * NT;
      if (idx < NP) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          if (kDir) buf[c * NP + idx] = preb[kDir ? s % kDepth : 0][k][c] + beta * pre[s % kDepth][k][c];
          else buf[c * NP + idx] = pre[s % kDepth][k][c];
        }
      }
    }
    if (s + kDepth <= ZC + 1) fetch(p + kDepth, s % kDepth);  // in flight for the next kDepth steps
    T e_cur = T(0);
    if (s >= 1) {
      e_cur = eq[(s - 1) & 1];  // layer p-1
      if (s + 2 <= ZC + 1) eq[(s - 1) & 1] = fetch_e(p + 1);
    }
    // the node finished in this step lies in plane L = p - 1 (steps s >= 2)
    const int L = p - 1;
    const bool outp = (s >= 2) && (L <= nz);
    Epi E;
    if (s >= 2) {
      if constexpr (kEpiAhead) {
        E = ep[s & 1];
        if (s + 1 <= ZC + 1) fetch_epi(L + 1, ep[(s + 1) & 1]);
      } else {
        fetch_epi(L, E);
      }
    }
    const uint32_t lb = plane_off(L);  // byte offset of the output plane

    __syncthreads();
    T Fcur[3][4];
    {
      const int o = tx + LX * ty;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const T* sc = buf + c * NP;
        face_wht(sc[o], sc[o + 1], sc[o + LX], sc[o + LX + 1], Fcur[c]);
      }
    }
    if (s >= 1) {
      // element layer L between planes p-1 and p
      // no branch on Ee == 0: lanes of one wave would run both paths, and
      // the scaled coefficients make an empty element contribute zeros anyway
      T gb[3][4], gt[3][4];
      {
        T S[3][4], D[3][4];
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
          for (int m = 0; m < 4; ++m) {
            S[c][m] = Fcur[c][m] + Fprev[c][m];
            D[c][m] = Fcur[c][m] - Fprev[c][m];
          }
        element_walsh(S, D, scaled(a.k, e_cur), gb, gt);
      }
      if (outp) {
        // plane L has both of its element layers: finish it
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          T G[4];
#pragma unroll
          for (int m = 0; m < 4; ++m) G[m] = gup[c][m] + gb[c][m];
          T c00, c10, c01, c11;
          face_iwht(G, c00, c10, c01, c11);
          sp[4 * c + 0][tid] = c00;
          sp[4 * c + 1][tid] = c10;
          sp[4 * c + 2][tid] = c01;
          sp[4 * c + 3][tid] = c11;
        }
        __syncthreads();
        {
          const T* prevbuf = &su[(s + 2) % 3][0][0];  // plane L
          const int o = (tx + 1) + LX * (ty + 1);
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            // element (tx,ty) -> its (1,1) corner, (tx+1,ty) -> (0,1), (tx,ty+1) -> (1,0), (tx+1,ty+1) -> (0,0)
            const T v = sp[4 * c + 3][tid] + sp[4 * c + 2][tid + 1] + sp[4 * c + 1][tid + TX] +
                        sp[4 * c + 0][tid + TX + 1];
            const T uc = prevbuf[c * NP + o];
            T out;
            if (MODE == kApply || MODE == kApplyDir) out = v;
            else if (MODE == kResid) out = E.bv[c] - v;
            else if (MODE == kJacobi) out = uc + a.omega * T(E.dv) * (E.bv[c] - v);
            else if (MODE == kCheb0) out = uc + a.c1 * uc + a.c2 * T(E.dv) * (E.bv[c] - v);
            else {
              // previous iterate: loaded, or x1 = omega0 D^-1 b of a smoothing that started from zero
              const T pv = a.prev_j0 ? a.omega0 * T(E.dv) * E.bv[c] : E.pv[c];
              out = uc + a.c1 * (uc - pv) + a.c2 * T(E.dv) * (E.bv[c] - v);
            }
            if (!((E.mk >> c) & 1)) out = T(0);
            bstore<T>(out, ry, outp ? ovo : kOob, lb + c * cstep);  // dropped for lanes without a node
            if (MODE == kApplyDir) bstore<T>(uc, rpo, outp ? ovo : kOob, lb + c * cstep);  // the direction
            if (MODE == kApply || MODE == kApplyDir) {
              if (a.dot_part && ncol) dot += (double)uc * (double)out;
            } else if (MODE == kCheb || MODE == kJacobi) {
              // <b, y>: the PCG's <r, z> when this step ends the fine-level V-cycle
              if (a.dot_part && ncol) dot += (double)E.bv[c] * (double)out;
            } else if (MODE == kResid) {
              // <y, y>: the squared norm of the (true) residual
              if (a.dot_part && ncol) dot += (double)out * (double)out;
            }
          }
        }
      }
#pragma unroll
      for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int m = 0; m < 4; ++m) gup[c][m] = gt[c][m];
    }
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int m = 0; m < 4; ++m) Fprev[c][m] = Fcur[c][m];
  }
  if ((MODE == kApply || MODE == kApplyDir || MODE == kCheb || MODE == kJacobi || MODE == kResid) && a.dot_part) {
    // fixed-order block reduction
    for (int off = 32; off > 0; off >>= 1) dot += __shfl_down(dot, off, 64);
    if ((tid & 63) == 0) sred[tid >> 6] = dot;
    __syncthreads();
    if (tid == 0) {
      double sum = 0.0;
      for (int w = 0; w < NT / 64; ++w) sum += sred[w];
      a.dot_part[(int64_t)r * nb + lid] = sum;
    }
  }
}

// ---------------------------------------------------------------------------
// Jacobi diagonal: diag(n) = kdiag * sum of Ee over the <= 8 elements at n.
// One launch serves every multigrid level: workgroups are assigned to levels
// by the (workgroup-aligned) node offsets in first[].
__global__ __launch_bounds__(256) void k_inv_diag(DiagLevels lv) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  int l = 0;
  while (l + 1 < lv.nlev && t >= lv.first[l + 1]) ++l;
  const float* __restrict__ Ee = lv.Ee[l];
  const uint8_t* __restrict__ mask = lv.mask[l];
  float* __restrict__ dinv = lv.dinv[l];
  const int nx = lv.nx[l], ny = lv.ny[l], nz = lv.nz[l];
  const double kdiag = lv.kdiag[l];
  const int64_t nn = (int64_t)(nx + 1) * (ny + 1) * (nz + 1);
  const int64_t n = t - lv.first[l];
  if (n >= nn) return;
  // 32-bit index arithmetic (a level has fewer than 2^31 nodes)
  const uint32_t n32 = uint32_t(n), sx = uint32_t(nx + 1), q = n32 / sx;
  const int i = int(n32 - q * sx), j = int(q % uint32_t(ny + 1)), k = int(q / uint32_t(ny + 1));
  // the 8 elements at the node, loads first (clamped in-grid addresses), summed in (dk, dj, di) order
  float ev[8];
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const int ei = i - 1 + (c & 1), ej = j - 1 + ((c >> 1) & 1), ek = k - 1 + (c >> 2);
    const bool in = ei >= 0 && ei < nx && ej >= 0 && ej < ny && ek >= 0 && ek < nz;
    const uint32_t e = uint32_t(min(max(ei, 0), nx - 1)) +
                       uint32_t(nx) * (uint32_t(min(max(ej, 0), ny - 1)) + uint32_t(ny) * uint32_t(min(max(ek, 0), nz - 1)));
    ev[c] = in ? Ee[e] : 0.f;
  }
  double s = 0.0;
#pragma unroll
  for (int c = 0; c < 8; ++c) s += ev[c];
  const double d = s * kdiag;
  dinv[n] = (mask[n] != 0 && d > 0.0) ? float(1.0 / d) : 0.f;
}

// Rediscretised coarse stiffness scale: mean of the (up to 8) children.
// Coarse-element stiffness scale from its (up to) 8 children: the power mean
// (mean of E^q)^(1/q); q = 1 is the arithmetic mean (an upper bound of the
// block's stiffness), smaller q lean toward the soft children, which matters
// where solid and void voxels mix.  Missing children (odd sizes) count as 0.
__global__ void k_coarsen_Ee(const float* __restrict__ Ef, float* __restrict__ Ec, int nx, int ny, int nz, int cx,
                             int cy, int cz, float q) {
  const int64_t nc = (int64_t)cx * cy * cz;
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= nc) return;
  const uint32_t e32 = uint32_t(e), qx = e32 / uint32_t(cx);
  const int I = int(e32 - qx * uint32_t(cx)), J = int(qx % uint32_t(cy)), K = int(qx / uint32_t(cy));
  // the 8 children, loads first (clamped in-grid addresses), summed in (dk, dj, di) order
  float v[8];
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const int i = 2 * I + (c & 1), j = 2 * J + ((c >> 1) & 1), k = 2 * K + (c >> 2);
    const bool in = i < nx && j < ny && k < nz;
    v[c] = Ef[uint32_t(min(i, nx - 1)) + uint32_t(nx) * (uint32_t(min(j, ny - 1)) + uint32_t(ny) * uint32_t(min(k, nz - 1)))];
    if (!in) v[c] = -1.f;  // missing child
  }
  double s = 0.0;
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    if (v[c] < 0.f) continue;
    const double d = v[c];
    // q = 1/2 (the default) as sqrt and a square: the same power mean without pow
    s += q == 1.0f ? d : (q == 0.5f ? sqrt(d) : pow(d, (double)q));
  }
  const double m = s * 0.125;
  Ec[e] = float(q == 1.0f ? m : (q == 0.5f ? m * m : pow(m, 1.0 / (double)q)));
}

// 1-D stencil of the trilinear prolongation seen from coarse index I: fine
// nodes 2I-1, 2I, 2I+1 with weights 1/2, 1, 1/2 (zero outside [0, n]);
// out-of-range indices are clamped so every load stays inside the grid.
__device__ __forceinline__ void fw_stencil(int I, int n, int (&idx)[3], float (&w)[3]) {
#pragma unroll
  for (int t = 0; t < 3; ++t) {
    const int i = 2 * I - 1 + t;
    const bool ok = i >= 0 && i <= n;
    idx[t] = ok ? i : 0;
    w[t] = ok ? (t == 1 ? 1.0f : 0.5f) : 0.0f;
  }
}

// Full-weighting restriction b_c = P^T r (P = trilinear prolongation): one
// thread per (coarse node, component), 27 loads each, fixed summation order.
// (A thread per node with all three components made the coarse levels
// latency-bound: too few waves, 81 loads each.)
template <typename T>
__global__ void k_restrict(const T* __restrict__ rf, T* __restrict__ bc, const uint8_t* __restrict__ maskc, int nx,
                           int ny, int nz, int cx, int cy, int cz, int64_t ldf, int64_t ldc, int nrhs,
                           const OpEntry* __restrict__ ops, int rpo, const double* __restrict__ skip) {
  const int64_t nc = (int64_t)(cx + 1) * (cy + 1) * (cz + 1);
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= 3 * nc) return;
  const int c = int(t / nc);
  const int64_t N = t - (int64_t)c * nc;
  const int r = blockIdx.y;
  if (rhs_skipped(skip, r)) return;
  if (ops) maskc = ops[r / rpo].mask;
  const int I = N % (cx + 1), J = (N / (cx + 1)) % (cy + 1), K = N / ((int64_t)(cx + 1) * (cy + 1));
  const uint8_t mk = maskc[N];
  int ix[3], iy[3], iz[3];
  float wx[3], wy[3], wz[3];
  fw_stencil(I, nx, ix, wx);
  fw_stencil(J, ny, iy, wy);
  fw_stencil(K, nz, iz, wz);
  const int64_t sx = nx + 1, sxy = (int64_t)(nx + 1) * (ny + 1);
  const T* src = rf + (int64_t)(3 * r + c) * ldf;
  float acc = 0.f;
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      const float w = wz[a] * wy[b];
      const T* row = src + iz[a] * sxy + iy[b] * sx;
      acc += w * (wx[0] * (float)row[ix[0]] + wx[1] * (float)row[ix[1]] + wx[2] * (float)row[ix[2]]);
    }
  bc[(int64_t)(3 * r + c) * ldc + N] = ((mk >> c) & 1) ? T(acc) : T(0);
}

// Prolongation with correction: x_f += M_f P x_c (up to 8 coarse nodes per fine node).
template <typename T>
__global__ void k_prolong_add(const T* __restrict__ xc, T* __restrict__ xf, const uint8_t* __restrict__ maskf, int nx,
                              int ny, int nz, int cx, int cy, int cz, int64_t ldf, int64_t ldc, int nrhs,
                              const OpEntry* __restrict__ ops, int rpo, const double* __restrict__ skip) {
  const int64_t nf = (int64_t)(nx + 1) * (ny + 1) * (nz + 1);
  const int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (n >= nf) return;
  const int r = blockIdx.y;
  if (rhs_skipped(skip, r)) return;
  if (ops) maskf = ops[r / rpo].mask;
  const uint8_t mk = maskf[n];
  if (!mk) return;
  const int i = n % (nx + 1), j = (n / (nx + 1)) % (ny + 1), k = n / ((int64_t)(nx + 1) * (ny + 1));
  // coarse neighbours I0 = i/2 and I1 = I0 + (i odd) per axis; with weight 1/2
  // on each of the pair (which coincide for even i) the 8-corner sum times
  // 1/8 is exactly the trilinear interpolation
  const int I0 = i >> 1, J0 = j >> 1, K0 = k >> 1;
  const int I1 = I0 + (i & 1), J1 = J0 + (j & 1), K1 = K0 + (k & 1);
  const int64_t csx = cx + 1, csxy = (int64_t)(cx + 1) * (cy + 1);
  const int64_t o00 = I0 + csx * J0 + csxy * K0, o01 = I1 + csx * J0 + csxy * K0;
  const int64_t o10 = I0 + csx * J1 + csxy * K0, o11 = I1 + csx * J1 + csxy * K0;
  const int64_t p00 = I0 + csx * J0 + csxy * K1, p01 = I1 + csx * J0 + csxy * K1;
  const int64_t p10 = I0 + csx * J1 + csxy * K1, p11 = I1 + csx * J1 + csxy * K1;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    if (!((mk >> c) & 1)) continue;
    const T* s = xc + (int64_t)(3 * r + c) * ldc;
    const float v = (((float)s[o00] + (float)s[o01]) + ((float)s[o10] + (float)s[o11])) +
                    (((float)s[p00] + (float)s[p01]) + ((float)s[p10] + (float)s[p11]));
    xf[(int64_t)(3 * r + c) * ldf + n] += T(0.125f * v);
  }
}

// ---------------------------------------------------------------------------
// Host-side launchers.  Tile shapes (TX x TY element columns per workgroup,
// ZC node planes per z chunk); the FP32 operator picks one per grid with a
// timing sweep (FemContext::autotune), the FP64 operator uses the first.
template <int TX, int TY, int ZC>
struct Tile {
  static constexpr int tx = TX, ty = TY, zc = ZC;
};
#define ANNEAL_TILES(X) X(0, 16, 16, 8) X(1, 16, 16, 16) X(2, 32, 8, 8) X(3, 32, 8, 16) X(4, 32, 16, 8) \
  X(5, 16, 16, 4) X(6, 64, 4, 16) X(7, 16, 16, 2) X(8, 8, 8, 2) X(9, 8, 8, 4) X(10, 16, 8, 2)

static void tile_dims(int variant, int& TX, int& TY, int& ZC) {
  switch (variant) {
#define ANNEAL_TILE_CASE(id, a, b, c) \
  case id:                           \
    TX = a;                          \
    TY = b;                          \
    ZC = c;                          \
    return;
    ANNEAL_TILES(ANNEAL_TILE_CASE)
#undef ANNEAL_TILE_CASE
    default: throw std::invalid_argument("unknown apply tile variant");
  }
}

int apply_variants() { return 11; }

int apply_blocks(int nx, int ny, int nz, int variant) {
  int TX, TY, ZC;
  tile_dims(variant, TX, TY, ZC);
  return ((nx + 1 + TX - 2) / (TX - 1)) * ((ny + 1 + TY - 2) / (TY - 1)) * ((nz + 1 + ZC - 1) / ZC);
}

template <typename T, int MODE, int TX, int TY, int ZC>
static void launch_tile(const ApplyArgs<T>& a0, int nrhs, hipStream_t s) {
  ApplyArgs<T> a = a0;
  ANNEAL_CHECK(3 * a.ldn * (int64_t)sizeof(T) < (int64_t(1) << 31) &&
                   (int64_t)a.nx * a.ny * a.nz * (int64_t)sizeof(T) < (int64_t(1) << 31),
               "grid too large for the 31-bit array offsets of the apply kernel");
  a.ntx = (a.nx + 1 + TX - 2) / (TX - 1);
  a.nty = (a.ny + 1 + TY - 2) / (TY - 1);
  a.nzc = (a.nz + 1 + ZC - 1) / ZC;
  dim3 grid(a.ntx * a.nty * a.nzc, nrhs);
  hipLaunchKernelGGL((k_apply<T, TX, TY, ZC, MODE>), grid, dim3(TX * TY), 0, s, a);
  ANNEAL_HIP(hipGetLastError());
}

template <typename T, int MODE>
static void launch_apply_mode(const ApplyArgs<T>& a, int nrhs, int variant, hipStream_t s) {
  if (sizeof(T) == 8) variant = kApplyVariant64;  // FP64: one shape (register pressure)
  switch (variant) {
#define ANNEAL_TILE_LAUNCH(id, a_, b_, c_) \
  case id:                                 \
    launch_tile<T, MODE, a_, b_, c_>(a, nrhs, s); \
    return;
    ANNEAL_TILES(ANNEAL_TILE_LAUNCH)
#undef ANNEAL_TILE_LAUNCH
    default: throw std::invalid_argument("unknown apply tile variant");
  }
}

template <typename T>
void launch_apply(const ApplyArgs<T>& a, int mode, int nrhs, hipStream_t s) {
  const int v = a.variant;
  switch (mode) {
    case kApply: launch_apply_mode<T, kApply>(a, nrhs, v, s); break;
    case kResid: launch_apply_mode<T, kResid>(a, nrhs, v, s); break;
    case kJacobi: launch_apply_mode<T, kJacobi>(a, nrhs, v, s); break;
    case kCheb: launch_apply_mode<T, kCheb>(a, nrhs, v, s); break;
    case kCheb0: launch_apply_mode<T, kCheb0>(a, nrhs, v, s); break;
    case kApplyDir:
      if constexpr (sizeof(T) == 4) {
        ANNEAL_CHECK(a.pout != nullptr && a.b != nullptr && a.skip != nullptr && (const void*)a.pout != (const void*)a.u,
                     "the direction apply needs z (b), the PCG scalars (skip) and a separate output direction");
        launch_apply_mode<T, kApplyDir>(a, nrhs, v, s);
        break;
      }
      throw std::invalid_argument("the direction apply is FP32 only");
    default: throw std::invalid_argument("bad apply mode");
  }
}
template void launch_apply<float>(const ApplyArgs<float>&, int, int, hipStream_t);
template void launch_apply<double>(const ApplyArgs<double>&, int, int, hipStream_t);

static inline unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

void launch_inv_diag(DiagLevels d, hipStream_t s) {
  ANNEAL_CHECK(d.nlev >= 1 && d.nlev <= DiagLevels::kMax, "bad level count for the Jacobi diagonals");
  // level l's nodes start at a multiple of the workgroup size
  d.first[0] = 0;
  for (int l = 0; l < d.nlev; ++l) {
    const int64_t nn = (int64_t)(d.nx[l] + 1) * (d.ny[l] + 1) * (d.nz[l] + 1);
    d.first[l + 1] = d.first[l] + (nn + 255) / 256 * 256;
  }
  hipLaunchKernelGGL(k_inv_diag, dim3(nblk(d.first[d.nlev], 256)), dim3(256), 0, s, d);
  ANNEAL_HIP(hipGetLastError());
}
void launch_coarsen_Ee(const float* Ef, float* Ec, const GridDims& f, const GridDims& c, float q, hipStream_t s) {
  hipLaunchKernelGGL(k_coarsen_Ee, dim3(nblk(c.nelem(), 256)), dim3(256), 0, s, Ef, Ec, f.nx, f.ny, f.nz, c.nx, c.ny,
                     c.nz, q);
  ANNEAL_HIP(hipGetLastError());
}
template <typename T>
void launch_restrict(const T* rf, T* bc, const uint8_t* maskc, const GridDims& f, const GridDims& c, int nrhs,
                     hipStream_t s, const OpEntry* ops, int rpo, const double* skip) {
  hipLaunchKernelGGL(k_restrict<T>, dim3(nblk(3 * c.nnode(), 256), nrhs), dim3(256), 0, s, rf, bc, maskc, f.nx, f.ny, f.nz,
                     c.nx, c.ny, c.nz, f.ldn(), c.ldn(), nrhs, ops, rpo, skip);
  ANNEAL_HIP(hipGetLastError());
}
template <typename T>
void launch_prolong_add(const T* xc, T* xf, const uint8_t* maskf, const GridDims& f, const GridDims& c, int nrhs,
                        hipStream_t s, const OpEntry* ops, int rpo, const double* skip) {
  hipLaunchKernelGGL(k_prolong_add<T>, dim3(nblk(f.nnode(), 256), nrhs), dim3(256), 0, s, xc, xf, maskf, f.nx, f.ny, f.nz,
                     c.nx, c.ny, c.nz, f.ldn(), c.ldn(), nrhs, ops, rpo, skip);
  ANNEAL_HIP(hipGetLastError());
}
template void launch_restrict<float>(const float*, float*, const uint8_t*, const GridDims&, const GridDims&, int,
                                     hipStream_t, const OpEntry*, int, const double*);
template void launch_prolong_add<float>(const float*, float*, const uint8_t*, const GridDims&, const GridDims&, int,
                                        hipStream_t, const OpEntry*, int, const double*);

}  // namespace anneal
