// Declarations shared by the solver, stencil and vector kernels.
#pragma once
#include "common.hpp"

namespace anneal {

constexpr int kDotBlocks = 512;  // fixed reduction grid: results do not depend on timing

// Per right-hand-side PCG scalars on the device (FP64).
constexpr int kSt = 8;
// kTol2: the inner tolerance on |r|^2 of a rhs; the PCG step that reaches it sets kDone on the device
constexpr int kRZ = 0, kRZnew = 1, kPQ = 2, kTol2 = 3, kAlpha = 4, kBeta = 5, kDone = 6, kZQ = 7;

// Kernels of a PCG inner loop take an optional `skip` pointer (the per-rhs
// scalar block above): the workgroups of a right-hand side whose done flag is
// set exit at entry and write nothing, so a converged rhs costs no bandwidth
// while the others of its block keep iterating.  Null means every rhs runs.
__device__ __forceinline__ bool rhs_skipped(const double* skip, int r) {
  return skip != nullptr && skip[r * kSt + kDone] != 0.0;
}

template <typename T>
struct KeCoef;
template <typename T>
struct ApplyArgs;

template <typename T>
void launch_apply(const ApplyArgs<T>& a, int mode, int nrhs, hipStream_t s);
int apply_blocks(int nx, int ny, int nz, int variant);
constexpr int kApplyVariant64 = 0;  // the one tile shape of the FP64 apply (register pressure)
int apply_variants();

// Jacobi diagonals of up to kMax multigrid levels in one launch (first[] is
// filled by the launcher).
struct DiagLevels {
  static constexpr int kMax = 8;
  const float* Ee[kMax];
  const uint8_t* mask[kMax];
  float* dinv[kMax];
  int nx[kMax], ny[kMax], nz[kMax];
  double kdiag[kMax];
  int64_t first[kMax + 1];
  int nlev;
};
void launch_inv_diag(DiagLevels d, hipStream_t s);
void launch_coarsen_Ee(const float* Ef, float* Ec, const GridDims& f, const GridDims& c, float q, hipStream_t s);
// With ops set (a device table of the level the mask belongs to), rhs r uses
// ops[r / rpo].mask instead of the mask argument.
template <typename T>
void launch_restrict(const T* rf, T* bc, const uint8_t* maskc, const GridDims& f, const GridDims& c, int nrhs,
                     hipStream_t s, const OpEntry* ops = nullptr, int rpo = 1, const double* skip = nullptr);
template <typename T>
void launch_prolong_add(const T* xc, T* xf, const uint8_t* maskf, const GridDims& f, const GridDims& c, int nrhs,
                        hipStream_t s, const OpEntry* ops = nullptr, int rpo = 1, const double* skip = nullptr);

// Reductions into out[r] (device); out_host, when set, is a device alias of
// pinned host memory that receives the same values (read after a wait on s).
template <typename TA, typename TB>
void dot(const TA* a, const TB* b, int64_t n, int64_t ld, int nrhs, double* part, double* out, hipStream_t s,
         double* out_host = nullptr);
void finalize(const double* part, int nblk, int nrhs, double* out, hipStream_t s, double* out_host = nullptr);
void jacobi0(const float* b, float* x, const float* dinv, const uint8_t* mask, float omega, int64_t n, int64_t ld,
             int nrhs, hipStream_t s, const double* skip = nullptr);
// With ops set, rhs r uses ops[r / rpo].Ainv.
void dense_solve(const double* Ainv, const float* b, float* x, int nnode, int64_t ld, int nrhs, hipStream_t s,
                 const OpEntry* ops = nullptr, int rpo = 1, const double* skip = nullptr);
void dense_assemble_invert(const float* Ee, const uint8_t* mask, const double* ke, double* A, double* B, int nx,
                           int ny, int nz, hipStream_t s);
// st[field] = sum of the partials, then op: 0 none; 1 alpha = rz / t; 2 beta = t / rz, rz = t;
// 3 (flexible) beta = -t / pq with t = <z_new, K p>; 4 rz = t.
void finalize_st(const double* part, int nblk, int nrhs, double* st, int field, int op, hipStream_t s);
void dot_partial(const float* a, const float* b, int64_t n, int64_t ld, int nrhs, double* part, hipStream_t s,
                 const double* skip = nullptr);
void pcg_update(float* x, float* r, const float* p, const float* q, const double* st, int64_t n, int64_t ld, int nrhs,
                double* part, double* out_rr, hipStream_t s, const double* skip = nullptr,
                double* out_host = nullptr);
// pcg_update with the convergence test on the device: rr = |r|^2 per rhs goes to
// rr_host; a live rhs with rr <= st[kTol2] gets st[kDone] = 1, so the kernels
// queued after it skip that rhs; done_host receives st[kDone] (both pinned).
void pcg_update_converge(float* x, float* r, const float* p, const float* q, double* st, int64_t n, int64_t ld,
                         int nrhs, double* part, hipStream_t s, double* rr_host, double* done_host);
void pcg_dir(float* p, const float* z, const double* st, int64_t n, int64_t ld, int nrhs, int first, hipStream_t s);
void to_f32_scaled(const double* a, float* b, const double* sc, int64_t n, int64_t ld, int nrhs, hipStream_t s);
// PCG pass set-up on the device from |r|^2 (rr) and |F|^2 (ff) per rhs: a rhs
// with |r| <= rtol |F| is done (scale 0, done flag set), otherwise scale 1/|r|;
// the scalar blocks start from zero but for the inner tolerance kTol2 =
// max(factor rtol |F| / |r|, floor)^2 (1 for a done rhs).  The host forms the
// same pass decisions from the same norms after its next wait.
void pass_setup(const double* rr, const double* ff, double rtol, double factor, double floor_, double* scale,
                double* st, int nrhs, hipStream_t s);
// dst1[0, n1) = src1[0, n1) and dst2[0, n2) = src2[0, n2) in one small kernel;
// src1 / src2 are device aliases of pinned host memory (n2 = 0: no second copy)
void upload(const double* src1, double* dst1, int n1, const double* src2, double* dst2, int n2, hipStream_t s);
void axpy_f32_to_f64(const float* x, double* y, const double* sc, int64_t n, int64_t ld, int nrhs, hipStream_t s);
void cast_f64_f32(const double* a, float* b, int64_t n, hipStream_t s);
// a64 = a and b = (float)a in one pass
void copy_cast_f64(const double* a, double* a64, float* b, int64_t n, hipStream_t s);

}  // namespace anneal
