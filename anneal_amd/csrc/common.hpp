// Shared definitions for the anneal HIP library (gfx950 / CDNA4).
//
// Layout conventions used by every kernel:
//   * Grid of nx*ny*nz hex8 voxels, element e = i + nx*(j + ny*k) (x fastest),
//     node n = i + (nx+1)*(j + (ny+1)*k)  (reference: anneal/fem.py:87-127).
//   * Nodal vectors are SoA: component c of right-hand side r lives at
//     base + (3*r + c) * ldn, with ldn >= nnode (a multiple of 64).
//   * Element fields (Ee, densities, sensitivities) cover the full grid; an
//     inactive voxel carries Ee = 0.
//   * dof_mask[n] bit c set <=> dof 3n+c is free (touched by an active element
//     and not fixed), i.e. the reference's FEM.free set (fem.py:145-155).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <string>
#include <stdexcept>

namespace anneal {

// Thread-local last error for the C ABI.
void set_error(const std::string& msg);
const char* get_error();

struct HipError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

#define ANNEAL_HIP(call)                                                            \
  do {                                                                              \
    hipError_t _e = (call);                                                         \
    if (_e != hipSuccess) {                                                         \
      char _b[512];                                                                 \
      snprintf(_b, sizeof(_b), "%s failed at %s:%d: %s", #call, __FILE__, __LINE__, \
               hipGetErrorString(_e));                                              \
      throw ::anneal::HipError(_b);                                                 \
    }                                                                               \
  } while (0)

#define ANNEAL_CHECK(cond, msg)                       \
  do {                                                \
    if (!(cond)) throw std::invalid_argument(msg);    \
  } while (0)

inline int64_t round_up(int64_t v, int64_t m) { return (v + m - 1) / m * m; }

// Consecutive workgroups are dealt round-robin to the 8 XCDs of the MI355X;
// this remaps a hardware block index to a logical tile index so that each XCD
// receives a contiguous run of logical tiles, keeping the halos that
// neighbouring tiles share inside one L2.
__device__ __forceinline__ int xcd_logical(int bid, int nb) {
  const int q = nb >> 3, r = nb & 7, x = bid & 7;
  return x * q + min(x, r) + (bid >> 3);
}


// Operator data of one problem on one multigrid level, for kernels that serve
// several problems of the same grid shape in one launch (right-hand side r
// belongs to problem r / rpo).  A null table means one operator for all rhs.
struct OpEntry {
  const float* Ee;     // FP32 stiffness scales of the level
  const double* Ee64;  // FP64 stiffness scales (finest level)
  const uint8_t* mask; // free-dof bits per node
  const float* dinv;   // Jacobi diagonal inverse
  const double* Ainv;  // dense inverse (coarsest level)
};

struct GridDims {
  int nx, ny, nz;
  int64_t nelem() const { return (int64_t)nx * ny * nz; }
  int64_t nnode() const { return (int64_t)(nx + 1) * (ny + 1) * (nz + 1); }
  int64_t ldn() const { return round_up(nnode(), 64); }
};

}  // namespace anneal
