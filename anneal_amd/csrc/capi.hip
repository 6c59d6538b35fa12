// extern "C" boundary of libanneal_hip.so (declared in include/anneal_hip.h).
#include "../../include/anneal_hip.h"

#include <algorithm>
#include <cstring>
#include <exception>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "batch.hpp"
#include "fem_ctx.hpp"
#include "fields.hpp"
#include "topopt.hpp"
#include "unet.hpp"
#include "voxel.hpp"

namespace anneal {
static thread_local std::string g_err;
void set_error(const std::string& m) { g_err = m; }
const char* get_error() { return g_err.c_str(); }
}  // namespace anneal

using namespace anneal;

// Static scratch of anneal_dot_f64_noalloc: the per-block partial sums of one
// call (kDotBlocks per right-hand side, at most kNoAllocMaxRhs right-hand
// sides, 256 KiB), so that the call allocates nothing.  Each device holds its
// own copy at its own address until the process ends; the host only ever uses
// the address hipGetSymbolAddress reports for the current device (the
// variable's name is a host-side stand-in).  Calls that overlap on one device
// share the array and race.
constexpr int kNoAllocMaxRhs = 64;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) double scratch[kNoAllocMaxRhs * kDotBlocks];

namespace {
template <typename F>
int guarded(F&& f) {
  try {
    anneal::set_error("");
    return f();
  } catch (const std::invalid_argument& e) {
    set_error(e.what());
    return 1;
  } catch (const std::exception& e) {
    set_error(e.what());
    return 2;
  } catch (...) {
    set_error("unknown error");
    return 2;
  }
}
struct FemHandle {
  FemContext* fem;
  uint8_t* flags;  // device copy of the active set (bit0)
};
inline FemHandle* fem_of(void* p) {
  if (!p) throw std::invalid_argument("null FEM context");
  return static_cast<FemHandle*>(p);
}
inline UNet* net_of(void* p) {
  if (!p) throw std::invalid_argument("null U-Net context");
  return static_cast<UNet*>(p);
}
inline FilterContext* filt_of(void* p) {
  if (!p) throw std::invalid_argument("null filter context");
  return static_cast<FilterContext*>(p);
}

std::mutex g_scratch_mu;
double* g_scratch_dev[kMaxDevices] = {};

// The address of `scratch` on the current device, looked up on the first call
// on that device (with it current) and kept.
double* scratch_here() {
  int dev = 0;
  ANNEAL_HIP(hipGetDevice(&dev));
  ANNEAL_CHECK(dev >= 0 && dev < kMaxDevices, "device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr)
      throw std::runtime_error(std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
    g_scratch_dev[dev] = static_cast<double*>(p);
  }
  return g_scratch_dev[dev];
}
}  // namespace

extern "C" {

const char* anneal_last_error(void) { return get_error(); }
int anneal_abi_version(void) { return 1; }

int anneal_ke_host(double nu, double h, double* ke_host, double* coef_host) {
  return guarded([&] {
    ANNEAL_CHECK(ke_host && coef_host, "null argument");
    ANNEAL_CHECK(nu > -1.0 && nu < 0.5 && h > 0.0, "need -1 < nu < 0.5 and h > 0");
    hex8_stiffness(nu, h, ke_host);
    const KeCoef<double> k = walsh_coefficients(ke_host);
    const double c[8] = {k.nAmB, k.nB, k.g, k.pc, k.pd, k.tAmB, k.tB, k.w};
    for (int i = 0; i < 8; ++i) coef_host[i] = c[i];
    return 0;
  });
}

int anneal_fem_create(void** ctx, int nx, int ny, int nz, double h, double nu, const uint8_t* active_host,
                      const uint8_t* fixed_host, int nrhs_max, int max_coarse_nodes, void* stream) {
  return guarded([&] {
    ANNEAL_CHECK(ctx && active_host && fixed_host, "null argument");
    ANNEAL_CHECK(h > 0.0, "voxel size must be positive");
    const int64_t ne = (int64_t)nx * ny * nz;
    bool any = false;
    for (int64_t e = 0; e < ne && !any; ++e) any = active_host[e] != 0;
    if (!any) throw std::invalid_argument("No active voxels: the design space is empty.");
    ANNEAL_CHECK(max_coarse_nodes >= 0 && max_coarse_nodes <= 4096, "max_coarse_nodes must be in [0, 4096]");
    auto* fem = new FemContext(nx, ny, nz, h, nu, active_host, fixed_host, nrhs_max, (hipStream_t)stream,
                               max_coarse_nodes > 0 ? max_coarse_nodes : kDefaultCoarseNodes);
    uint8_t* flags = nullptr;
    try {
      ANNEAL_HIP(hipMalloc(&flags, ne));
      std::string tmp(ne, '\0');
      for (int64_t e = 0; e < ne; ++e) tmp[e] = active_host[e] ? kActive : 0;
      ANNEAL_HIP(hipMemcpy(flags, tmp.data(), ne, hipMemcpyHostToDevice));
    } catch (...) {
      delete fem;
      throw;
    }
    *ctx = new FemHandle{fem, flags};
    return 0;
  });
}

int anneal_fem_destroy(void* ctx) {
  return guarded([&] {
    if (!ctx) return 0;
    FemHandle* h = static_cast<FemHandle*>(ctx);
    delete h->fem;
    (void)hipFree(h->flags);
    delete h;
    return 0;
  });
}

int anneal_fem_info(void* ctx, int64_t* ldn, int* levels, uint8_t* mask_host) {
  return guarded([&] {
    FemHandle* h = fem_of(ctx);
    if (ldn) *ldn = h->fem->ldn();
    if (levels) *levels = h->fem->levels();
    if (mask_host) ANNEAL_HIP(hipMemcpy(mask_host, h->fem->mask(), h->fem->grid().nnode(), hipMemcpyDeviceToHost));
    return 0;
  });
}

int anneal_fem_level_info(void* ctx, int level, int* nx, int* ny, int* nz, int* tile) {
  return guarded([&] {
    FemContext* f = fem_of(ctx)->fem;
    ANNEAL_CHECK(level >= 0 && level < f->levels(), "level out of range");
    const Level& L = f->level(level);
    if (nx) *nx = L.g.nx;
    if (ny) *ny = L.g.ny;
    if (nz) *nz = L.g.nz;
    if (tile) *tile = L.tile;
    return 0;
  });
}

int anneal_fem_set_Ee(void* ctx, const double* Ee_dev) {
  return guarded([&] {
    ANNEAL_CHECK(Ee_dev, "null Ee");
    fem_of(ctx)->fem->set_Ee(Ee_dev);
    return 0;
  });
}

int anneal_fem_rebuild(void* ctx) {
  return guarded([&] {
    fem_of(ctx)->fem->rebuild_coarse();
    return 0;
  });
}

int anneal_fem_apply_f32(void* ctx, const float* u_dev, float* y_dev, int nrhs) {
  return guarded([&] {
    ANNEAL_CHECK(u_dev && y_dev && u_dev != (const float*)y_dev, "apply needs distinct u and y");
    ANNEAL_CHECK(nrhs >= 1 && nrhs <= 64, "nrhs out of range");
    fem_of(ctx)->fem->apply_f32(u_dev, y_dev, nrhs);
    return 0;
  });
}

int anneal_fem_apply_f64(void* ctx, const double* u_dev, double* y_dev, int nrhs) {
  return guarded([&] {
    ANNEAL_CHECK(u_dev && y_dev && u_dev != (const double*)y_dev, "apply needs distinct u and y");
    ANNEAL_CHECK(nrhs >= 1 && nrhs <= 64, "nrhs out of range");
    fem_of(ctx)->fem->apply_f64(u_dev, y_dev, nrhs);
    return 0;
  });
}

int anneal_fem_solve(void* ctx, const double* F_dev, double* U_dev, int nrhs, double rtol, int maxiter, int use_x0,
                     int* iters_host) {
  return guarded([&] {
    ANNEAL_CHECK(F_dev && U_dev && iters_host, "null argument");
    ANNEAL_CHECK(rtol > 0.0 && maxiter >= 1, "rtol must be positive and maxiter >= 1");
    const int st = fem_of(ctx)->fem->solve(F_dev, U_dev, nrhs, rtol, maxiter, use_x0 != 0, iters_host);
    if (st == 1) {
      set_error("not converged within maxiter");
      return 3;
    }
    if (st == 2) {
      set_error("non-finite residual");
      return 4;
    }
    return 0;
  });
}

int anneal_fem_autotune(void* ctx, int variant, int* chosen_host, float* us_host, int* n_host) {
  return guarded([&] {
    FemHandle* h = fem_of(ctx);
    const int c = variant == -2 ? h->fem->tile() : h->fem->autotune(variant, us_host);
    if (chosen_host) *chosen_host = c;
    if (n_host) *n_host = apply_variants();
    return 0;
  });
}

int anneal_fem_set_level_tile(void* ctx, int level, int variant) {
  return guarded([&] {
    fem_of(ctx)->fem->set_level_tile(level, variant);
    return 0;
  });
}

int anneal_fem_set_option(void* ctx, const char* name, double value) {
  return guarded([&] {
    ANNEAL_CHECK(name, "null option name");
    FemContext* f = fem_of(ctx)->fem;
    const std::string n(name);
    if (n == "graphs") {
      f->use_graphs = value != 0.0;
    } else if (n == "smoother_degree") {
      ANNEAL_CHECK(value >= 1 && value <= 15, "smoother_degree must be in [1, 15]");
      f->nu_pre = (int)value;
      f->invalidate_graphs();
    } else if (n == "smoother_degree_coarse") {
      ANNEAL_CHECK(value >= 1 && value <= 15, "smoother_degree_coarse must be in [1, 15]");
      f->nu_coarse = (int)value;
      f->invalidate_graphs();
    } else if (n == "smoother_degree_deep") {
      ANNEAL_CHECK(value >= 0 && value <= 15, "smoother_degree_deep must be in [0, 15] (0: as the coarse levels)");
      f->nu_deep = (int)value;
      f->invalidate_graphs();
    } else if (n == "deep_escalate_iters") {
      ANNEAL_CHECK(value >= 0, "deep_escalate_iters must be >= 0 (0: off)");
      f->deep_escalate_iters = (int)value;
    } else if (n == "deep_escalate_degree") {
      ANNEAL_CHECK(value >= 1 && value <= 15, "deep_escalate_degree must be in [1, 15]");
      f->deep_escalate_degree = (int)value;
    } else if (n == "coarsen_power") {
      ANNEAL_CHECK(value > 0.0 && value <= 1.0, "coarsen_power must be in (0, 1]");
      f->coarsen_power = value;
    } else if (n == "inner_tol_factor") {
      ANNEAL_CHECK(value > 0.0 && value < 1.0, "inner_tol_factor must be in (0, 1)");
      f->inner_tol_factor = value;
    } else if (n == "cheb_ratio") {
      ANNEAL_CHECK(value > 1.0, "cheb_ratio must exceed 1");
      f->cheb_ratio = value;
      f->invalidate_graphs();
    } else if (n == "inner_tol_floor") {
      ANNEAL_CHECK(value > 0.0 && value < 1.0, "inner_tol_floor must be in (0, 1)");
      f->inner_tol_floor = value;
    } else if (n == "fuse_first_step") {
      f->fuse_first_from = value != 0.0 ? 0 : 99;
      f->invalidate_graphs();
    } else if (n == "fuse_first_level") {
      ANNEAL_CHECK(value >= 0.0, "fuse_first_level must be >= 0");
      f->fuse_first_from = (int)value;
      f->invalidate_graphs();
    } else if (n == "rebuild_each_solve") {
      f->rebuild_each_solve = value != 0.0;
    } else if (n == "inner_max") {
      ANNEAL_CHECK(value >= 1, "inner_max must be >= 1");
      f->inner_max = (int)value;
    } else {
      throw std::invalid_argument("unknown option: " + n);
    }
    return 0;
  });
}

int anneal_fem_smooth_step_f32(void* ctx, const float* u_dev, const float* b_dev, float* y_dev, int nrhs) {
  return guarded([&] {
    ANNEAL_CHECK(u_dev && b_dev && y_dev && (const float*)y_dev != u_dev && (const float*)y_dev != b_dev,
                 "smoothing step needs u, b and a distinct in-place y");
    FemHandle* h = fem_of(ctx);
    ANNEAL_CHECK(h->flags, "smoothing steps run on the problems of a group");
    ANNEAL_CHECK(nrhs >= 1 && nrhs <= 64, "nrhs out of range");
    h->fem->smooth_step_f32(u_dev, b_dev, y_dev, nrhs);
    return 0;
  });
}

int anneal_fem_vcycle(void* ctx, const float* r_dev, float* z_dev, int nrhs) {
  return guarded([&] {
    ANNEAL_CHECK(r_dev && z_dev && r_dev != (const float*)z_dev, "vcycle needs distinct r and z");
    fem_of(ctx)->fem->set_rhs_count(nrhs);
    fem_of(ctx)->fem->vcycle(0, r_dev, z_dev, nrhs);
    return 0;
  });
}

int anneal_fem_element_energy(void* ctx, const double* U_dev, int nrhs, double* ce_dev) {
  return guarded([&] {
    ANNEAL_CHECK(U_dev && ce_dev, "null argument");
    FemHandle* h = fem_of(ctx);
    ANNEAL_CHECK(h->flags, "element energies are computed on the problems of a group");
    element_energy(U_dev, nrhs, h->fem->ldn(), h->fem->grid(), h->fem->coef64(), h->flags, ce_dev,
                   h->fem->stream());
    return 0;
  });
}

int anneal_fem_stress(void* ctx, const double* U_dev, int nrhs, const double* B0_host, const double* D_host,
                      const int* local_host, const double* E_dev, double E, double* sig_dev, double* vm_dev) {
  return guarded([&] {
    ANNEAL_CHECK(U_dev && B0_host && D_host && local_host && sig_dev && nrhs >= 1, "null argument");
    FemHandle* h = fem_of(ctx);
    ANNEAL_CHECK(h->flags, "stresses are computed on the problems of a group");
    StressMats m;
    for (int a = 0; a < 6; ++a) {
      for (int b = 0; b < 24; ++b) m.B0[a][b] = B0_host[a * 24 + b];
      for (int b = 0; b < 6; ++b) m.D[a][b] = D_host[a * 6 + b];
    }
    for (int q = 0; q < 8; ++q)
      for (int c = 0; c < 3; ++c) {
        m.local[q][c] = local_host[q * 3 + c];
        ANNEAL_CHECK(m.local[q][c] == 0 || m.local[q][c] == 1, "corner offsets must be 0 or 1");
      }
    element_stress(U_dev, nrhs, h->fem->ldn(), h->fem->grid(), m, h->flags, E_dev, E, sig_dev, vm_dev,
                   h->fem->stream());
    return 0;
  });
}

int anneal_fem_sensitivity(void* ctx, const double* U_dev, int nrhs, const uint8_t* flags_dev, const double* xphys_dev,
                           const double* dxp_dev, const double* phi_dev, const double* coef_dev, double p, double emin,
                           double E0, double* dcdx_dev, double* dvdx_dev) {
  return guarded([&] {
    ANNEAL_CHECK(U_dev && flags_dev && xphys_dev && dxp_dev && phi_dev && coef_dev && dcdx_dev && dvdx_dev,
                 "null argument");
    FemHandle* h = fem_of(ctx);
    sensitivity(U_dev, nrhs, h->fem->ldn(), h->fem->grid(), h->fem->coef64(), flags_dev, xphys_dev, dxp_dev, phi_dev,
                coef_dev, p, emin, E0, dcdx_dev, dvdx_dev, h->fem->stream());
    return 0;
  });
}

int anneal_filter_create(void** ctx, int nx, int ny, int nz, double r_vox, const uint8_t* flags_host, void* stream) {
  return guarded([&] {
    ANNEAL_CHECK(ctx && flags_host, "null argument");
    *ctx = new FilterContext(nx, ny, nz, r_vox, flags_host, (hipStream_t)stream);
    return 0;
  });
}

int anneal_filter_destroy(void* ctx) {
  return guarded([&] {
    delete static_cast<FilterContext*>(ctx);
    return 0;
  });
}

int anneal_filter_forward(void* ctx, const double* x_dev, double* y_dev) {
  return guarded([&] {
    ANNEAL_CHECK(x_dev && y_dev && x_dev != (const double*)y_dev, "filter needs distinct input and output");
    filt_of(ctx)->forward(x_dev, y_dev);
    return 0;
  });
}

int anneal_filter_backward(void* ctx, const double* d_dev, double* y_dev) {
  return guarded([&] {
    ANNEAL_CHECK(d_dev && y_dev && d_dev != (const double*)y_dev, "filter needs distinct input and output");
    filt_of(ctx)->backward(d_dev, y_dev);
    return 0;
  });
}

int anneal_filter_backward_pair(void* ctx, const double* d1_dev, double* y1_dev, const double* d2_dev,
                                double* y2_dev, double* stats_dev) {
  return guarded([&] {
    ANNEAL_CHECK(d1_dev && y1_dev && d2_dev && y2_dev, "null argument");
    ANNEAL_CHECK(d1_dev != (const double*)y1_dev && d2_dev != (const double*)y2_dev && y1_dev != y2_dev &&
                     d1_dev != (const double*)y2_dev && d2_dev != (const double*)y1_dev,
                 "filter needs distinct inputs and outputs");
    filt_of(ctx)->backward_pair(d1_dev, y1_dev, d2_dev, y2_dev, stats_dev);
    return 0;
  });
}

int anneal_filter_physical(void* ctx, const double* x_dev, double beta, const double* phi_dev, double* xphys_dev,
                           double* dxp_dev, double* vol_dev) {
  return guarded([&] {
    ANNEAL_CHECK(x_dev && phi_dev && xphys_dev && dxp_dev, "null argument");
    filt_of(ctx)->physical(x_dev, beta, phi_dev, xphys_dev, dxp_dev, vol_dev);
    return 0;
  });
}

int anneal_oc_bisect(void* ctx, const double* x_dev, const double* dcx_dev, const double* dvx_dev,
                     const double* phi_dev, double beta, double move, double lo, double hi, double vtarget,
                     double hint, double hint_width, double* xnew_dev, int* steps_host, int* evals_host,
                     double* lam_last_host) {
  return guarded([&] {
    ANNEAL_CHECK(x_dev && dcx_dev && dvx_dev && phi_dev && xnew_dev, "null argument");
    ANNEAL_CHECK(lo > 0.0 && hi >= lo, "bisection bracket must satisfy 0 < lo <= hi");
    int ev = 0;
    double ll = 0.0;
    const int s = filt_of(ctx)->oc_bisect(x_dev, dcx_dev, dvx_dev, phi_dev, beta, move, lo, hi, vtarget, xnew_dev,
                                          hint, hint_width, &ev, &ll);
    if (steps_host) *steps_host = s;
    if (evals_host) *evals_host = ev;
    if (lam_last_host) *lam_last_host = ll;
    return 0;
  });
}

int anneal_oc_begin(void* ctx, const double* x_dev, const double* dcx_dev, const double* dvx_dev,
                    const double* phi_dev, double beta, double move, double lo, double hi, double vtarget, double hint,
                    double hint_width, double* xnew_dev) {
  return guarded([&] {
    ANNEAL_CHECK(x_dev && dcx_dev && dvx_dev && phi_dev && xnew_dev, "null argument");
    ANNEAL_CHECK(lo > 0.0 && hi >= lo, "bisection bracket must satisfy 0 < lo <= hi");
    filt_of(ctx)->oc_begin(x_dev, dcx_dev, dvx_dev, phi_dev, beta, move, lo, hi, vtarget, xnew_dev, hint, hint_width);
    return 0;
  });
}

int anneal_oc_begin_stats(void* ctx, const double* x_dev, const double* dcx_dev, const double* dvx_dev,
                          const double* phi_dev, double beta, double move, const double* stats_dev,
                          double hint_ratio, double vtarget, double hint_width, double* xnew_dev) {
  return guarded([&] {
    ANNEAL_CHECK(x_dev && dcx_dev && dvx_dev && phi_dev && stats_dev && xnew_dev, "null argument");
    filt_of(ctx)->oc_begin_stats(x_dev, dcx_dev, dvx_dev, phi_dev, beta, move, stats_dev, hint_ratio, vtarget,
                                 xnew_dev, hint_width);
    return 0;
  });
}

int anneal_oc_poll(void* ctx, int* done_host) {
  return guarded([&] {
    ANNEAL_CHECK(done_host, "null argument");
    *done_host = filt_of(ctx)->oc_poll() ? 1 : 0;
    return 0;
  });
}

int anneal_oc_end(void* ctx, int* steps_host, int* evals_host, double* lam_last_host) {
  return guarded([&] {
    int ev = 0;
    double ll = 0.0;
    const int s = filt_of(ctx)->oc_end(&ev, &ll);
    if (steps_host) *steps_host = s;
    if (evals_host) *evals_host = ev;
    if (lam_last_host) *lam_last_host = ll;
    return 0;
  });
}

int anneal_stiffness_law(const double* xphys_dev, const double* phi_dev, const uint8_t* flags_dev, double E0,
                         double emin, double p, int64_t n, double* Ee_dev, void* stream) {
  return guarded([&] {
    ANNEAL_CHECK(xphys_dev && phi_dev && flags_dev && Ee_dev, "null argument");
    if (n > 0) stiffness_law(xphys_dev, phi_dev, flags_dev, E0, emin, p, n, Ee_dev, (hipStream_t)stream);
    return 0;
  });
}

int anneal_design_stats(const double* xold_dev, const double* xnew_dev, const double* xphys_dev,
                        const double* phi_dev, const double* dcx_dev, const double* dvx_dev,
                        const uint8_t* flags_dev, int64_t n, double* out_dev, double* work_dev, void* stream) {
  return guarded([&] {
    ANNEAL_CHECK(phi_dev && flags_dev && out_dev && work_dev, "null argument");
    design_stats(xold_dev, xnew_dev, xphys_dev, phi_dev, dcx_dev, dvx_dev, flags_dev, n, out_dev, work_dev,
                 (hipStream_t)stream);
    return 0;
  });
}

int anneal_dot_f64(const double* a_dev, const double* b_dev, int64_t nnode, int64_t ldn, int nrhs, double* out_dev,
                   double* work_dev, void* stream) {
  return guarded([&] {
    ANNEAL_CHECK(a_dev && b_dev && out_dev && work_dev && nrhs >= 1, "null argument");
    dot(a_dev, b_dev, nnode, ldn, nrhs, work_dev, out_dev, (hipStream_t)stream);
    return 0;
  });
}

int anneal_dot_f64_noalloc(const double* a_dev, const double* b_dev, int64_t nnode, int64_t ldn, int nrhs,
                           double* out_dev, void* stream) {
  return guarded([&] {
    ANNEAL_CHECK(a_dev && b_dev && out_dev, "null argument");
    ANNEAL_CHECK(nrhs >= 1 && nrhs <= kNoAllocMaxRhs,
                 "anneal_dot_f64_noalloc: nrhs must be in [1, 64] (the size of its static scratch)");
    ANNEAL_CHECK(nnode >= 0 && ldn >= nnode, "anneal_dot_f64_noalloc: need 0 <= nnode <= ldn");
    // every partial the reduction reads is written by this call's first kernel, so
    // the scratch needs no clearing
    dot(a_dev, b_dev, nnode, ldn, nrhs, scratch_here(), out_dev, (hipStream_t)stream);
    return 0;
  });
}

int anneal_fields_corner_values(const float* u_dev, const float* vm_dev, int nx, int ny, int nz, double h,
                                const double* origin_host, const double* tri_dev, const double* nrm_dev,
                                int64_t nface, float* out_dev, void* stream) {
  return guarded([&] {
    ANNEAL_CHECK(nx >= 1 && ny >= 1 && nz >= 1 && h > 0.0, "need a non-empty grid and h > 0");
    ANNEAL_CHECK(nface >= 0, "nface must be >= 0");
    ANNEAL_CHECK(origin_host, "null origin");
    ANNEAL_CHECK(nface == 0 || (u_dev && vm_dev && tri_dev && nrm_dev && out_dev), "null argument");
    corner_values(u_dev, vm_dev, nx, ny, nz, h, origin_host, tri_dev, nrm_dev, nface, out_dev,
                  (hipStream_t)stream);
    return 0;
  });
}

int anneal_unet_create(void** ctx, int cin, int cout, const int* widths, int levels, void* stream) {
  return guarded([&] {
    ANNEAL_CHECK(ctx && widths && levels >= 2 && levels <= 6, "bad U-Net arguments");
    *ctx = nullptr;
    std::vector<int> w(widths, widths + levels);
    *ctx = new UNet(cin, cout, w, (hipStream_t)stream);
    return 0;
  });
}

int anneal_unet_destroy(void* ctx) {
  return guarded([&] {
    delete static_cast<UNet*>(ctx);
    return 0;
  });
}

int anneal_unet_nparams(void* ctx, int64_t* n) {
  return guarded([&] {
    ANNEAL_CHECK(n, "null argument");
    *n = net_of(ctx)->nparams();
    return 0;
  });
}

int anneal_unet_get_params(void* ctx, float* host) {
  return guarded([&] {
    ANNEAL_CHECK(host, "null argument");
    net_of(ctx)->get_params(host);
    return 0;
  });
}

int anneal_unet_set_params(void* ctx, const float* host) {
  return guarded([&] {
    ANNEAL_CHECK(host, "null argument");
    net_of(ctx)->set_params(host);
    return 0;
  });
}

int anneal_unet_get_grads(void* ctx, float* host) {
  return guarded([&] {
    ANNEAL_CHECK(host, "null argument");
    net_of(ctx)->get_grads(host);
    return 0;
  });
}

int anneal_unet_forward(void* ctx, const uint16_t* in_dev, int B, int nx, int ny, int nz, float* out_dev) {
  return guarded([&] {
    ANNEAL_CHECK(in_dev && out_dev, "null argument");
    ANNEAL_CHECK(B >= 1 && nx >= 1 && ny >= 1 && nz >= 1, "empty input");
    net_of(ctx)->forward(in_dev, NetDims{B, nx, ny, nz}, out_dev);
    return 0;
  });
}

int anneal_unet_backward(void* ctx, const float* dout_dev) {
  return guarded([&] {
    ANNEAL_CHECK(dout_dev, "null argument");
    net_of(ctx)->backward(dout_dev);
    return 0;
  });
}

int anneal_unet_zero_grad(void* ctx) {
  return guarded([&] {
    net_of(ctx)->zero_grad();
    return 0;
  });
}

int anneal_unet_adam(void* ctx, float lr, float beta1, float beta2, float eps, int step, float weight_decay) {
  return guarded([&] {
    net_of(ctx)->adam(lr, beta1, beta2, eps, step, weight_decay);
    return 0;
  });
}

int anneal_unet_rel_mse(const float* o_dev, const float* t_dev, int B, int nx, int ny, int nz, int c, float* dout_dev,
                        double* loss_dev, double* work_dev, void* stream) {
  return guarded([&] {
    ANNEAL_CHECK(o_dev && t_dev && loss_dev && work_dev, "null argument");
    ANNEAL_CHECK(B >= 1 && B <= 64 && c >= 1, "need 1 <= B <= 64 samples");
    net_rel_mse(o_dev, t_dev, NetDims{B, nx, ny, nz}, c, dout_dev, loss_dev, work_dev, (hipStream_t)stream);
    return 0;
  });
}

int anneal_unet_restrict(const uint16_t* x_dev, int B, int fx, int fy, int fz, int c, uint16_t* y_dev, float s,
                         void* stream) {
  return guarded([&] {
    ANNEAL_CHECK(x_dev && y_dev && c % 8 == 0 && c > 0, "bad restrict arguments");
    const NetDims f{B, fx, fy, fz}, co{B, coarse_n(fx), coarse_n(fy), coarse_n(fz)};
    net_restrict(View{(bf16*)x_dev, c, c}, f, View{y_dev, c, c}, co, s, nullptr, false, (hipStream_t)stream);
    return 0;
  });
}

int anneal_unet_prolong(const uint16_t* x_dev, int B, int cx, int cy, int cz, int c, uint16_t* y_dev, int fx, int fy,
                        int fz, float s, int accumulate, void* stream) {
  return guarded([&] {
    ANNEAL_CHECK(x_dev && y_dev && c % 8 == 0 && c > 0, "bad prolong arguments");
    ANNEAL_CHECK(cx == coarse_n(fx) && cy == coarse_n(fy) && cz == coarse_n(fz), "coarse dims must be n/2 + 1");
    net_prolong(View{(bf16*)x_dev, c, c}, NetDims{B, cx, cy, cz}, View{y_dev, c, c}, NetDims{B, fx, fy, fz}, s,
                nullptr, accumulate != 0, (hipStream_t)stream);
    return 0;
  });
}

int anneal_fem_set_learned(void* fem, void* unet, int level, double e_ref, int mode) {
  return guarded([&] {
    FemHandle* h = fem_of(fem);
    h->fem->set_learned(unet ? net_of(unet) : nullptr, level, e_ref, mode);
    h->fem->invalidate_graphs();
    return 0;
  });
}

int anneal_fem_learned_stage(void* fem, const float* r_dev, float* z_dev, int nrhs) {
  return guarded([&] {
    ANNEAL_CHECK(r_dev && z_dev && nrhs >= 1, "null argument");
    FemHandle* h = fem_of(fem);
    ANNEAL_CHECK(h->fem->learned_mode() != 0, "no network attached (anneal_fem_set_learned)");
    h->fem->learned_stage(r_dev, z_dev, nrhs);
    return 0;
  });
}

int anneal_fem_learned_pair(void* fem, const double* F_dev, const double* U_dev, int nrhs, int level, double e_ref,
                            uint16_t* in_dev, float* tgt_dev) {
  return guarded([&] {
    ANNEAL_CHECK(F_dev && U_dev && in_dev && tgt_dev, "null argument");
    fem_of(fem)->fem->learned_pair(F_dev, U_dev, nrhs, level, e_ref, in_dev, tgt_dev);
    return 0;
  });
}

int anneal_fem_group_create(void** ctx, void* const* members, int nmember, int nrhs_max, void* stream) {
  return guarded([&] {
    ANNEAL_CHECK(ctx && members && nmember >= 1 && nmember <= 64, "need 1..64 member FEM contexts");
    *ctx = nullptr;
    std::vector<FemContext*> v;
    for (int i = 0; i < nmember; ++i) v.push_back(fem_of(members[i])->fem);
    auto* fem = new FemContext(v, nrhs_max, (hipStream_t)stream);
    *ctx = new FemHandle{fem, nullptr};
    return 0;
  });
}

int anneal_batch_create(void** batch, void* const* fems, int nprob) {
  return guarded([&] {
    ANNEAL_CHECK(batch && fems && nprob >= 1 && nprob <= 64, "need 1..64 FEM contexts");
    *batch = nullptr;
    std::vector<FemContext*> v;
    for (int i = 0; i < nprob; ++i) v.push_back(fem_of(fems[i])->fem);
    *batch = new BatchSolver(v);
    return 0;
  });
}

int anneal_batch_destroy(void* batch) {
  return guarded([&] {
    delete static_cast<BatchSolver*>(batch);
    return 0;
  });
}

int anneal_batch_solve(void* batch, const double* const* F_dev, double* const* U_dev, const int* nrhs, double rtol,
                       int maxiter, int use_x0, int* status, int* iters) {
  return guarded([&] {
    ANNEAL_CHECK(batch && F_dev && U_dev && nrhs && status && iters, "null argument");
    ANNEAL_CHECK(rtol > 0.0 && maxiter >= 1, "need rtol > 0 and maxiter >= 1");
    BatchSolver* b = static_cast<BatchSolver*>(batch);
    const int np = b->size();
    std::vector<const double*> F(F_dev, F_dev + np);
    std::vector<double*> U(U_dev, U_dev + np);
    std::vector<int> nr(nrhs, nrhs + np);
    for (int p = 0; p < np; ++p) ANNEAL_CHECK(F[p] && U[p], "null F or U");
    std::vector<int> st(np);
    b->solve(F, U, nr, rtol, maxiter, use_x0 != 0, st.data(), iters);
    int worst = 0;
    for (int p = 0; p < np; ++p) {
      status[p] = st[p] == 0 ? 0 : (st[p] == 1 ? 3 : 4);
      worst = std::max(worst, status[p]);
    }
    if (worst) set_error("some problems of the batch did not converge (see status)");
    return worst;
  });
}


// ---- voxel passes (problem.py, export.py) -----------------------------------
namespace {
GridDims voxel_grid(int nx, int ny, int nz) {
  ANNEAL_CHECK(nx >= 1 && ny >= 1 && nz >= 1, "need a non-empty grid");
  return GridDims{nx, ny, nz};
}
}  // namespace

int anneal_region_select(int nx, int ny, int nz, double h, const double* origin_host, const int* kind_dev,
                         const double* prm_dev, const int* first_dev, const double* thr_dev, int nanchor,
                         int at_nodes, const uint8_t* emask_dev, uint32_t* bits_dev, long long* counts_dev,
                         void* stream) {
  return guarded([&] {
    const GridDims g = voxel_grid(nx, ny, nz);
    ANNEAL_CHECK(h > 0.0 && origin_host, "need h > 0 and an origin");
    ANNEAL_CHECK(nanchor >= 0 && nanchor <= kMaxAnchors, "nanchor must be in [0, 32]");
    ANNEAL_CHECK(emask_dev && bits_dev && counts_dev && (nanchor == 0 || (kind_dev && prm_dev && first_dev && thr_dev)),
                 "null argument");
    region_select(g, h, origin_host, RegionTable{kind_dev, prm_dev, first_dev, thr_dev, nanchor}, at_nodes != 0,
                  emask_dev, bits_dev, counts_dev, (hipStream_t)stream);
    return 0;
  });
}

int anneal_region_sdf(int64_t npts, const double* pts_dev, const int* kind_dev, const double* prm_dev,
                      const int* first_dev, int nanchor, double* out_dev, void* stream) {
  return guarded([&] {
    ANNEAL_CHECK(npts >= 0 && nanchor >= 0 && nanchor <= kMaxAnchors, "bad sizes");
    ANNEAL_CHECK(npts == 0 || nanchor == 0 || (pts_dev && kind_dev && prm_dev && first_dev && out_dev),
                 "null argument");
    region_sdf(npts, pts_dev, RegionTable{kind_dev, prm_dev, first_dev, nullptr, nanchor}, out_dev,
               (hipStream_t)stream);
    return 0;
  });
}

int anneal_boundary_fields(int nx, int ny, int nz, const uint32_t* node_bits_dev, int nanchor,
                           const uint8_t* dofmask_dev, const double* load_dev, int ncase, uint8_t* fixed_bits_dev,
                           double* F_dev, void* stream) {
  return guarded([&] {
    const GridDims g = voxel_grid(nx, ny, nz);
    ANNEAL_CHECK(nanchor >= 0 && nanchor <= kMaxAnchors && ncase >= 0, "bad sizes");
    ANNEAL_CHECK(node_bits_dev && fixed_bits_dev && (nanchor == 0 || dofmask_dev) &&
                     (ncase == 0 || (F_dev && (nanchor == 0 || load_dev))),
                 "null argument");
    boundary_fields(g, node_bits_dev, nanchor, dofmask_dev, load_dev, ncase, fixed_bits_dev, F_dev,
                    (hipStream_t)stream);
    return 0;
  });
}

int anneal_gather_node_bits(int nx, int ny, int nz, const uint32_t* node_bits_dev, const uint8_t* emask_dev,
                            uint32_t* ebits_dev, void* stream) {
  return guarded([&] {
    const GridDims g = voxel_grid(nx, ny, nz);
    ANNEAL_CHECK(node_bits_dev && emask_dev && ebits_dev, "null argument");
    gather_node_bits(g, node_bits_dev, emask_dev, ebits_dev, (hipStream_t)stream);
    return 0;
  });
}

int anneal_voxel_components(int nx, int ny, int nz, const uint8_t* mask_dev, int32_t* label_dev,
                            long long* work_dev, int64_t* ncomp_out, void* stream) {
  return guarded([&] {
    const GridDims g = voxel_grid(nx, ny, nz);
    ANNEAL_CHECK(mask_dev && label_dev && work_dev && ncomp_out, "null argument");
    *ncomp_out = components(g, mask_dev, label_dev, work_dev, (hipStream_t)stream);
    return 0;
  });
}

int anneal_voxel_link(int nx, int ny, int nz, const uint8_t* mask_dev, const int32_t* label_dev,
                      const uint32_t* ebits_dev, uint32_t seed, uint8_t* keep_dev, uint8_t* out_mask_dev,
                      long long* work_dev, int64_t* nkept_out, uint32_t* linked_out, void* stream) {
  return guarded([&] {
    const GridDims g = voxel_grid(nx, ny, nz);
    ANNEAL_CHECK(mask_dev && label_dev && ebits_dev && keep_dev && work_dev && nkept_out, "null argument");
    *nkept_out = link_components(g, mask_dev, label_dev, ebits_dev, seed, keep_dev, out_mask_dev, linked_out,
                                 work_dev, (hipStream_t)stream);
    return 0;
  });
}

int anneal_voxel_threshold(int64_t n, const double* f_dev, double t, uint8_t* mask_dev, void* stream) {
  return guarded([&] {
    ANNEAL_CHECK(n >= 0 && (n == 0 || (f_dev && mask_dev)), "bad arguments");
    threshold_mask(n, f_dev, t, mask_dev, (hipStream_t)stream);
    return 0;
  });
}

int anneal_voxel_count_above(int64_t n, const double* f_dev, double t, long long* work_dev, int64_t* count_out,
                             void* stream) {
  return guarded([&] {
    ANNEAL_CHECK(n >= 0 && work_dev && count_out && (n == 0 || f_dev), "bad arguments");
    *count_out = count_above(n, f_dev, t, work_dev, (hipStream_t)stream);
    return 0;
  });
}

int anneal_output_field(int64_t n, const double* x_dev, const double* fill_dev, const uint8_t* active_dev,
                        const double* solid_fill_dev, const double* void_fill_dev, const uint8_t* frozen_dev,
                        double* f_dev, void* stream) {
  return guarded([&] {
    ANNEAL_CHECK(n >= 0, "n must be >= 0");
    ANNEAL_CHECK(n == 0 || (x_dev && fill_dev && active_dev && solid_fill_dev && void_fill_dev && frozen_dev && f_dev),
                 "null argument");
    output_field(n, x_dev, fill_dev, active_dev, solid_fill_dev, void_fill_dev, frozen_dev, f_dev,
                 (hipStream_t)stream);
    return 0;
  });
}

int64_t anneal_choose_level_work_bytes(void) { return (int64_t)choose_level_work_bytes(); }

int anneal_choose_level(int64_t n, const double* f_dev, const double* x_dev, const double* fill_dev, void* work_dev,
                        double* level_out, double* target_out, void* stream) {
  return guarded([&] {
    ANNEAL_CHECK(n >= 0 && work_dev && level_out && (n == 0 || (f_dev && x_dev && fill_dev)), "bad arguments");
    *level_out = choose_level(n, f_dev, x_dev, fill_dev, work_dev, target_out, (hipStream_t)stream);
    return 0;
  });
}

int anneal_paths_sssp(int nx, int ny, int nz, const uint8_t* active_dev, const double* f_dev, double need,
                      const uint32_t* ebits_dev, uint32_t source_groups, double* dist_dev, int32_t* pred_dev,
                      int32_t* stamp_dev, int* flag_dev, int* sweeps_out, void* stream) {
  return guarded([&] {
    const GridDims g = voxel_grid(nx, ny, nz);
    ANNEAL_CHECK(active_dev && f_dev && ebits_dev && dist_dev && pred_dev && stamp_dev && flag_dev, "null argument");
    ANNEAL_CHECK(std::isfinite(need), "need must be finite");
    const int bound = (int)std::min<int64_t>(g.nelem() + 2, (int64_t)1 << 30);
    const int sw = sssp(g, active_dev, f_dev, need, ebits_dev, source_groups, dist_dev, pred_dev, stamp_dev, flag_dev,
                        bound, (hipStream_t)stream);
    if (sweeps_out) *sweeps_out = sw;
    return 0;
  });
}

int anneal_paths_trace(int nx, int ny, int nz, const uint8_t* active_dev, const uint32_t* ebits_dev, int group,
                       const double* dist_dev, const int32_t* pred_dev, int32_t* path_dev, long long* work_dev,
                       int64_t* start_out, int64_t* len_out, double* dstart_out, void* stream) {
  return guarded([&] {
    const GridDims g = voxel_grid(nx, ny, nz);
    ANNEAL_CHECK(active_dev && ebits_dev && dist_dev && pred_dev && path_dev && work_dev && start_out && len_out,
                 "null argument");
    *len_out = trace_path(g, active_dev, ebits_dev, group, dist_dev, pred_dev, path_dev, work_dev, start_out,
                          dstart_out, (hipStream_t)stream);
    return 0;
  });
}

int anneal_paths_raise(int nx, int ny, int nz, const int32_t* path_dev, int64_t len, const uint8_t* active_dev,
                       const double* fill_dev, double need, double* f_dev, void* stream) {
  return guarded([&] {
    const GridDims g = voxel_grid(nx, ny, nz);
    ANNEAL_CHECK(len >= 0 && len <= g.nelem(), "path length out of range");
    ANNEAL_CHECK(len == 0 || (path_dev && active_dev && fill_dev && f_dev), "null argument");
    raise_path(g, path_dev, len, active_dev, fill_dev, need, f_dev, (hipStream_t)stream);
    return 0;
  });
}

}  // extern "C"
