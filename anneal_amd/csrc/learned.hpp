// Network input/output on a multigrid level's node grid (see learned.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "common.hpp"

namespace anneal {

constexpr int kNetCin = 8, kNetCout = 3;

// out[b][n][8] bf16 from the level's Ee, dof mask and residual r ([nrhs][3][ldn]);
// nrm2[b] = |r_b|^2 (device).
void net_pack(const float* Ee, const uint8_t* mask, const float* r, const double* nrm2, const GridDims& g, int nrhs,
              double e_ref, uint16_t* out, hipStream_t s);
// z[b][c][n] = mask ? o[b][n][c] * |r_b| / sqrt(V) / e_ref : 0 (whole ldn written).
void net_unpack(const float* o, const uint8_t* mask, const double* nrm2, const GridDims& g, int nrhs, double e_ref,
                float* z, hipStream_t s);
// Training target: the fine FP64 solution injected onto level `level` in network units.
void net_target(const double* U, const GridDims& fine, int level, const GridDims& g, const double* nrm2, int nrhs,
                double e_ref, float* t, hipStream_t s);
// gamma[b] = rz[b] / zkz[b] (energy-optimal step along the network correction; 0 if zkz <= 0)
void net_gamma(const double* rz, const double* zkz, int nrhs, double* gamma, hipStream_t s);
// t = r - gamma K z1 ; z1 *= gamma   ([nrhs][3][ld])
void net_scaled_resid(const float* r, const float* kz, float* z1, float* t, const double* gamma, int64_t ld, int nrhs,
                      hipStream_t s);
void fill_zero(float* x, int64_t n, hipStream_t s);
void add_f32(float* y, const float* x, int64_t n, hipStream_t s);

}  // namespace anneal
