// Sampling of saved result fields at display-mesh triangle corners.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace anneal {

// out[3*nface][4] = [dx, dy, dz, vonMises] per triangle corner
// (anneal/fields.py:17-37).  u: [nnode][3] float32 node field; vm: [nelem]
// float32 element field; tri: [nface][3][3] corner points; nrm: [nface][3]
// face normals; origin: 3 host doubles.  All arrays are device pointers.
void corner_values(const float* u, const float* vm, int nx, int ny, int nz, double h, const double* origin,
                   const double* tri, const double* nrm, int64_t nface, float* out, hipStream_t s);

}  // namespace anneal
