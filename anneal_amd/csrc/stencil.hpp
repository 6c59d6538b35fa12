// Matrix-free operator argument blocks shared by the apply kernels and the solver.
#pragma once
#include "common.hpp"
#include "blas.hpp"

namespace anneal {

template <typename T>
struct KeCoef {
  T nAmB, nB, g, pc, pd, tAmB, tB, w;
};

enum ApplyMode : int {
  kApply = 0,   // y = M K u
  kResid = 1,   // y = M (b - K u)
  kJacobi = 2,  // y = M (u + omega * dinv * (b - K u))
  kCheb = 3,    // y = M (u + c1 (u - prev) + c2 * dinv * (b - K u)); prev may be null (0) or alias y
  kCheb0 = 4,   // kCheb with u = M omega dinv b formed on the fly and prev = 0 (first two smoothing steps fused)
  kApplyDir = 5,  // the PCG direction and K times it: p = b + beta u (beta from the PCG scalar block) staged
                  // on the fly and stored to pout at the owned nodes, y = M K p (FP32 only)
};

template <typename T>
struct ApplyArgs {
  const T* u;          // [nrhs][3][ldn]
  T* y;                // [nrhs][3][ldn]
  const T* b;          // kResid / kJacobi
  const T* Ee;         // [nelem] (FP64 operator keeps FP64 stiffness scales)
  const uint8_t* mask; // [nnode]
  const float* dinv;   // [nnode] (kJacobi)
  double* dot_part;    // optional per-block partials [nrhs][nblk_xyz]: kApply sum u.y, kCheb / kJacobi sum b.y,
                       // kResid sum y.y
  KeCoef<T> k;
  T omega;
  const T* prev;       // kCheb: previous iterate (read at the node only, so it may alias y)
  T c1, c2;            // kCheb recurrence coefficients
  T omega0;            // kCheb with prev_j0: the previous iterate is omega0 D^-1 b (formed at the node)
  int prev_j0;
  int nx, ny, nz;
  int64_t ldn;
  int ntx, nty, nzc;   // tiles in x, y and z chunks (set by the launcher)
  int variant;         // tile shape id (host side)
  const OpEntry* ops;  // per-problem operators (device table) or null: Ee, mask, dinv above serve every rhs
  int rpo;             // right-hand sides per problem when ops is set
  const double* skip;  // PCG scalar block: rhs with the done flag set are skipped (null: none)
  T* pout;             // kApplyDir: the new direction [nrhs][3][ldn] (not u: other tiles still read u)
};

// Face (2-D) Walsh transform of 4 corner values a=(i,j) b=(i+1,j) c=(i,j+1) d=(i+1,j+1)
// into modes [1, xi, eta, xi*eta].
template <typename T>
__host__ __device__ __forceinline__ void face_wht(T a, T b, T c, T d, T* f) {
  T p0 = a + b, q0 = b - a, p1 = c + d, q1 = d - c;
  f[0] = p0 + p1;
  f[1] = q0 + q1;
  f[2] = p1 - p0;
  f[3] = q1 - q0;
}

// Inverse of face_wht (un-normalised): nodal values of modes g at the 4 corners.
template <typename T>
__host__ __device__ __forceinline__ void face_iwht(const T* g, T& a, T& b, T& c, T& d) {
  T s0 = g[0] - g[2], s1 = g[1] - g[3];  // row j   (s_b = -1)
  T t0 = g[0] + g[2], t1 = g[1] + g[3];  // row j+1 (s_b = +1)
  a = s0 - s1;
  b = s0 + s1;
  c = t0 - t1;
  d = t0 + t1;
}

// Element matvec in the Walsh basis.  S = top + bottom face modes (even in z),
// D = top - bottom (odd in z); indices [component][face mode].  YS / YD are the
// even / odd output modes (constant modes carry no energy and stay zero).
template <typename T>
__host__ __device__ __forceinline__ void element_walsh_modes(const T (&S)[3][4], const T (&D)[3][4],
                                                             const KeCoef<T>& k, T (&YS)[3][4], T (&YD)[3][4]) {
  // normal strains: (ux,xi) (uy,eta) (uz,zeta)
  {
    T s = S[0][1] + S[1][2] + D[2][0];
    T t = k.nB * s;
    YS[0][1] = k.nAmB * S[0][1] + t;
    YS[1][2] = k.nAmB * S[1][2] + t;
    YD[2][0] = k.nAmB * D[2][0] + t;
  }
  // shear pairs (rank one)
  {
    T q = k.g * (S[0][2] + S[1][1]);
    YS[0][2] = q;
    YS[1][1] = q;
    q = k.g * (D[0][0] + S[2][1]);
    YD[0][0] = q;
    YS[2][1] = q;
    q = k.g * (D[1][0] + S[2][2]);
    YD[1][0] = q;
    YS[2][2] = q;
  }
  // bilinear pairs
  YS[0][3] = k.pc * S[0][3] + k.pd * D[2][2];
  YD[2][2] = k.pd * S[0][3] + k.pc * D[2][2];
  YS[1][3] = k.pc * S[1][3] + k.pd * D[2][1];
  YD[2][1] = k.pd * S[1][3] + k.pc * D[2][1];
  YD[1][2] = k.pc * D[1][2] + k.pd * D[0][1];
  YD[0][1] = k.pd * D[1][2] + k.pc * D[0][1];
  // bilinear triple
  {
    T s = S[2][3] + D[0][2] + D[1][1];
    T t = k.tB * s;
    YS[2][3] = k.tAmB * S[2][3] + t;
    YD[0][2] = k.tAmB * D[0][2] + t;
    YD[1][1] = k.tAmB * D[1][1] + t;
  }
  // trilinear
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    YD[c][3] = k.w * D[c][3];
    YS[c][0] = T(0);
  }
}

// Face-domain outputs for the bottom (gb) and top (gt) node planes.
template <typename T>
__host__ __device__ __forceinline__ void element_walsh(const T (&S)[3][4], const T (&D)[3][4],
                                                       const KeCoef<T>& k, T (&gb)[3][4], T (&gt)[3][4]) {
  T YS[3][4], YD[3][4];
  element_walsh_modes(S, D, k, YS, YD);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    gb[c][0] = -YD[c][0];
    gt[c][0] = YD[c][0];
#pragma unroll
    for (int m = 1; m < 4; ++m) {
      gb[c][m] = YS[c][m] - YD[c][m];
      gt[c][m] = YS[c][m] + YD[c][m];
    }
  }
}

}  // namespace anneal
