// Multigrid hierarchy and mixed-precision PCG driver (host side of the solve).
#include "fem_ctx.hpp"

#include <cstring>
#include "learned.hpp"
#include "unet.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>

namespace anneal {

// ---------------------------------------------------------------------------
// Element stiffness (restates the quadrature of anneal/fem.py:36-84 with the
// corner order q = a + 2b + 4c).
void hex8_stiffness(double nu, double h, double* ke) {
  const double c = 1.0 / ((1.0 + nu) * (1.0 - 2.0 * nu));
  double D[6][6] = {};
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) D[a][b] = c * (a == b ? 1.0 - nu : nu);
  for (int a = 3; a < 6; ++a) D[a][a] = c * 0.5 * (1.0 - 2.0 * nu);
  const double gp = 1.0 / std::sqrt(3.0), half = 0.5 * h, w = half * half * half;
  std::fill(ke, ke + 576, 0.0);
  for (int iz = 0; iz < 2; ++iz)
    for (int iy = 0; iy < 2; ++iy)
      for (int ix = 0; ix < 2; ++ix) {
        const double xi = ix ? gp : -gp, eta = iy ? gp : -gp, zeta = iz ? gp : -gp;
        double B[6][24] = {};
        for (int q = 0; q < 8; ++q) {
          const double sa = (q & 1) ? 1.0 : -1.0, sb = (q & 2) ? 1.0 : -1.0, sc = (q & 4) ? 1.0 : -1.0;
          const double gx = sa * (1 + sb * eta) * (1 + sc * zeta) / 8.0 / half;
          const double gy = sb * (1 + sa * xi) * (1 + sc * zeta) / 8.0 / half;
          const double gz = sc * (1 + sa * xi) * (1 + sb * eta) / 8.0 / half;
          B[0][3 * q] = gx;
          B[1][3 * q + 1] = gy;
          B[2][3 * q + 2] = gz;
          B[3][3 * q] = gy;
          B[3][3 * q + 1] = gx;
          B[4][3 * q + 1] = gz;
          B[4][3 * q + 2] = gy;
          B[5][3 * q] = gz;
          B[5][3 * q + 2] = gx;
        }
        double DB[6][24];
        for (int a = 0; a < 6; ++a)
          for (int j = 0; j < 24; ++j) {
            double s = 0;
            for (int b = 0; b < 6; ++b) s += D[a][b] * B[b][j];
            DB[a][j] = s;
          }
        for (int i = 0; i < 24; ++i)
          for (int j = 0; j < 24; ++j) {
            double s = 0;
            for (int a = 0; a < 6; ++a) s += B[a][i] * DB[a][j];
            ke[i * 24 + j] += s * w;
          }
      }
}

static inline int widx(int comp, int fm, int mz) { return 3 * (fm + 4 * mz) + comp; }

KeCoef<double> walsh_coefficients(const double* ke) {
  // Khat = T^T ke T / 64 with T[3q+c][3m+c] = W[q][m], W = Walsh signs.
  double W[8][8];
  for (int q = 0; q < 8; ++q)
    for (int m = 0; m < 8; ++m) {
      const int a = q & 1, b = (q >> 1) & 1, c = (q >> 2) & 1;
      const int mx = m & 1, my = (m >> 1) & 1, mz = (m >> 2) & 1;
      double s = 1.0;
      if (mx && !a) s = -s;
      if (my && !b) s = -s;
      if (mz && !c) s = -s;
      W[q][m] = s;
    }
  double K[24][24];
  double maxabs = 0.0;
  for (int i = 0; i < 24; ++i)
    for (int j = 0; j < 24; ++j) {
      const int m = i / 3, ci = i % 3, m2 = j / 3, cj = j % 3;
      double s = 0.0;
      for (int q = 0; q < 8; ++q)
        for (int q2 = 0; q2 < 8; ++q2) s += W[q][m] * ke[(3 * q + ci) * 24 + 3 * q2 + cj] * W[q2][m2];
      K[i][j] = s / 64.0;
      maxabs = std::max(maxabs, std::fabs(K[i][j]));
    }
  KeCoef<double> k;
  const double nA = K[widx(0, 1, 0)][widx(0, 1, 0)];
  k.nB = K[widx(0, 1, 0)][widx(1, 2, 0)];
  k.nAmB = nA - k.nB;
  k.g = K[widx(0, 2, 0)][widx(0, 2, 0)];
  k.pc = K[widx(0, 3, 0)][widx(0, 3, 0)];
  k.pd = K[widx(0, 3, 0)][widx(2, 2, 1)];
  const double tA = K[widx(2, 3, 0)][widx(2, 3, 0)];
  k.tB = K[widx(2, 3, 0)][widx(0, 2, 1)];
  k.tAmB = tA - k.tB;
  k.w = K[widx(0, 3, 1)][widx(0, 3, 1)];
  // Verify: the structured matvec must reproduce every column of Khat.
  for (int j = 0; j < 24; ++j) {
    double S[3][4] = {}, Dm[3][4] = {};
    const int m = j / 3, c = j % 3, fm = m & 3, mz = m >> 2;
    (mz ? Dm : S)[c][fm] = 1.0;
    double gb[3][4], gt[3][4];
    element_walsh(S, Dm, k, gb, gt);
    for (int i = 0; i < 24; ++i) {
      const int mi = i / 3, cc = i % 3, fmi = mi & 3, mzi = mi >> 2;
      const double ys = 0.5 * (gb[cc][fmi] + gt[cc][fmi]), yd = 0.5 * (gt[cc][fmi] - gb[cc][fmi]);
      const double got = mzi ? yd : ys;
      const double want = K[i][j];
      if (std::fabs(got - want) > 1e-11 * maxabs) {
        char b[160];
        snprintf(b, sizeof b, "Walsh pattern of ke does not hold at (%d,%d): %g vs %g", i, j, got, want);
        throw std::runtime_error(b);
      }
    }
  }
  return k;
}

// lambda_max(ke) / ke_00 by power iteration (ke is symmetric positive semi-definite
// with a constant diagonal).
double element_jacobi_bound(const double* ke) {
  double v[24], w[24];
  for (int i = 0; i < 24; ++i) v[i] = 1.0 + 0.01 * i * (i % 3 == 0 ? 1.0 : -1.0);
  double lam = 0.0;
  for (int it = 0; it < 5000; ++it) {
    double nrm = 0.0;
    for (int i = 0; i < 24; ++i) {
      double s = 0.0;
      for (int j = 0; j < 24; ++j) s += ke[i * 24 + j] * v[j];
      w[i] = s;
      nrm += s * s;
    }
    nrm = std::sqrt(nrm);
    double vn = 0.0;
    for (int i = 0; i < 24; ++i) vn += v[i] * v[i];
    lam = nrm / std::sqrt(vn);
    for (int i = 0; i < 24; ++i) v[i] = w[i] / nrm;
  }
  return lam / ke[0];
}

template <typename T>
static KeCoef<T> cast_coef(const KeCoef<double>& k, double s) {
  return KeCoef<T>{T(k.nAmB * s), T(k.nB * s), T(k.g * s), T(k.pc * s), T(k.pd * s), T(k.tAmB * s), T(k.tB * s),
                   T(k.w * s)};
}

// ---------------------------------------------------------------------------
void* FemContext::alloc(size_t bytes) {
  void* p = nullptr;
  ANNEAL_HIP(hipMalloc(&p, std::max<size_t>(bytes, 256)));
  owned_.push_back(p);
  return p;
}

FemContext::FemContext(int nx, int ny, int nz, double h, double nu, const uint8_t* active_host,
                       const uint8_t* fixed_host, int nrhs_max, hipStream_t stream, int max_coarse_nodes)
    : s_(stream), nrhs_max_(nrhs_max) {
  ANNEAL_CHECK(nx >= 1 && ny >= 1 && nz >= 1, "grid dimensions must be positive");
  ANNEAL_CHECK(nrhs_max >= 1 && nrhs_max <= 64, "nrhs_max must be in [1, 64]");
  ANNEAL_CHECK(nu > -1.0 && nu < 0.5, "Poisson ratio must be in (-1, 0.5)");
  try {
    hex8_stiffness(nu, h, ke_);
    k64_ = walsh_coefficients(ke_);
    cheb_lmax = 1.001 * element_jacobi_bound(ke_);
    // Level geometry
    std::vector<GridDims> dims{GridDims{nx, ny, nz}};
    while (true) {
      const GridDims& g = dims.back();
      if (g.nnode() <= max_coarse_nodes || g.nx < 2 || g.ny < 2 || g.nz < 2) break;
      dims.push_back(GridDims{(g.nx + 1) / 2, (g.ny + 1) / 2, (g.nz + 1) / 2});
    }
    ANNEAL_CHECK(dims.back().nnode() <= 4096, "coarsest multigrid level too large for the dense solve");
    // Host masks per level
    std::vector<uint8_t> act(active_host, active_host + dims[0].nelem());
    std::vector<uint8_t> fix(fixed_host, fixed_host + dims[0].nnode());
    lv_.resize(dims.size());
    for (size_t l = 0; l < dims.size(); ++l) {
      const GridDims g = dims[l];
      if (l > 0) {
        const GridDims f = dims[l - 1];
        std::vector<uint8_t> a2(g.nelem(), 0), f2(g.nnode(), 0);
        for (int k = 0; k < f.nz; ++k)
          for (int j = 0; j < f.ny; ++j)
            for (int i = 0; i < f.nx; ++i)
              if (act[i + (int64_t)f.nx * (j + (int64_t)f.ny * k)])
                a2[(i / 2) + (int64_t)g.nx * ((j / 2) + (int64_t)g.ny * (k / 2))] = 1;
        for (int K = 0; K <= g.nz; ++K)
          for (int J = 0; J <= g.ny; ++J)
            for (int I = 0; I <= g.nx; ++I) {
              uint8_t bits = 0;
              for (int k = std::max(0, 2 * K - 1); k <= std::min(f.nz, 2 * K + 1); ++k)
                for (int j = std::max(0, 2 * J - 1); j <= std::min(f.ny, 2 * J + 1); ++j)
                  for (int i = std::max(0, 2 * I - 1); i <= std::min(f.nx, 2 * I + 1); ++i)
                    bits |= fix[i + (int64_t)(f.nx + 1) * (j + (int64_t)(f.ny + 1) * k)];
              f2[I + (int64_t)(g.nx + 1) * (J + (int64_t)(g.ny + 1) * K)] = bits;
            }
        act.swap(a2);
        fix.swap(f2);
      }
      std::vector<uint8_t> mk(g.nnode(), 0);
      for (int k = 0; k < g.nz; ++k)
        for (int j = 0; j < g.ny; ++j)
          for (int i = 0; i < g.nx; ++i)
            if (act[i + (int64_t)g.nx * (j + (int64_t)g.ny * k)])
              for (int c = 0; c < 8; ++c)
                mk[(i + (c & 1)) + (int64_t)(g.nx + 1) * ((j + ((c >> 1) & 1)) + (int64_t)(g.ny + 1) * (k + (c >> 2)))] =
                    7;
      for (int64_t n = 0; n < g.nnode(); ++n) mk[n] &= (uint8_t)(~fix[n] & 7);
      Level& L = lv_[l];
      L.g = g;
      L.h = h * double(1 << l);
      L.k = cast_coef<float>(k64_, double(1 << l));
      L.kdiag = ke_[0] * double(1 << l);
      L.mask = (uint8_t*)alloc(g.nnode());
      ANNEAL_HIP(hipMemcpy(L.mask, mk.data(), g.nnode(), hipMemcpyHostToDevice));
      L.Ee = (float*)alloc(g.nelem() * sizeof(float));
      L.dinv = (float*)alloc(g.nnode() * sizeof(float));
      const size_t vb = (size_t)nrhs_max * 3 * g.ldn() * sizeof(float);
      L.x = (float*)alloc(vb);
      L.x2 = (float*)alloc(vb);
      L.b = (float*)alloc(vb);
      L.r = (float*)alloc(vb);
      ANNEAL_HIP(hipMemset(L.x, 0, vb));
      ANNEAL_HIP(hipMemset(L.x2, 0, vb));
      ANNEAL_HIP(hipMemset(L.b, 0, vb));
      ANNEAL_HIP(hipMemset(L.r, 0, vb));
    }
    for (int i = 0; i < 24; ++i)
      if (std::fabs(ke_[i * 24 + i] - ke_[0]) > 1e-12 * std::fabs(ke_[0]))
        throw std::runtime_error("hex8 stiffness diagonal is not constant");
    const GridDims& gc = lv_.back().g;
    const int64_t nd = 3 * gc.nnode();
    Ainv_ = (double*)alloc(nd * nd * sizeof(double));
    Awork_ = (double*)alloc(nd * nd * sizeof(double));
    ke_coarse_dev_ = (double*)alloc(576 * sizeof(double));
    {
      double kc[576];
      const double sc = double(1 << (lv_.size() - 1));
      for (int i = 0; i < 576; ++i) kc[i] = ke_[i] * sc;
      ANNEAL_HIP(hipMemcpy(ke_coarse_dev_, kc, sizeof kc, hipMemcpyHostToDevice));
    }
    Ee64_ = (double*)alloc(lv_[0].g.nelem() * sizeof(double));
    alloc_solver(nrhs_max);
    // stiffness scales start at 0: the tile sweep below (and any apply before
    // the first operator is set) reads defined values
    for (auto& L : lv_) ANNEAL_HIP(hipMemset(L.Ee, 0, L.g.nelem() * sizeof(float)));
    if (lv_[0].g.nelem() >= 4096) pick_tiles();
  } catch (...) {
    for (void* p : owned_) (void)hipFree(p);
    owned_.clear();
    throw;
  }
}

void FemContext::alloc_solver(int nrhs_max) {
  const GridDims& g0 = lv_[0].g;
  const size_t v32 = (size_t)nrhs_max * 3 * g0.ldn() * sizeof(float);
  const size_t v64 = (size_t)nrhs_max * 3 * g0.ldn() * sizeof(double);
  px_ = (float*)alloc(v32);
  pr_ = (float*)alloc(v32);
  pz_ = (float*)alloc(v32);
  pp_ = (float*)alloc(v32);
  pp2_ = (float*)alloc(v32);
  pq_ = (float*)alloc(v32);
  x64_ = (double*)alloc(v64);
  r64_ = (double*)alloc(v64);
  ANNEAL_HIP(hipMemset(px_, 0, v32));
  ANNEAL_HIP(hipMemset(pr_, 0, v32));
  ANNEAL_HIP(hipMemset(pz_, 0, v32));
  ANNEAL_HIP(hipMemset(pp_, 0, v32));
  ANNEAL_HIP(hipMemset(pp2_, 0, v32));
  ANNEAL_HIP(hipMemset(pq_, 0, v32));
  ANNEAL_HIP(hipMemset(r64_, 0, v64));
  int nblk = kDotBlocks;
  for (int v = 0; v < apply_variants(); ++v) nblk = std::max(nblk, apply_blocks(g0.nx, g0.ny, g0.nz, v));
  part_ = (double*)alloc((size_t)nblk * nrhs_max * sizeof(double));
  st_ = (double*)alloc(64 * kSt * sizeof(double));
  red_ = (double*)alloc(4 * 64 * sizeof(double));
  scale_ = (double*)alloc(64 * sizeof(double));
  ANNEAL_HIP(hipHostMalloc((void**)&host_red_, 6 * 64 * sizeof(double), hipHostMallocDefault));
  ANNEAL_HIP(hipHostGetDevicePointer((void**)&host_red_dev_, host_red_, 0));
  ANNEAL_HIP(hipHostMalloc((void**)&host_up_, (64 + 64 * kSt) * sizeof(double), hipHostMallocDefault));
  ANNEAL_HIP(hipHostGetDevicePointer((void**)&host_up_dev_, host_up_, 0));
  ANNEAL_HIP(hipStreamCreateWithFlags(&cap_s_, hipStreamNonBlocking));
  ANNEAL_HIP(hipEventCreateWithFlags(&ev_, hipEventDisableTiming));
  ANNEAL_HIP(hipEventCreateWithFlags(&ev2_, hipEventDisableTiming));
  vgraph_.assign(nrhs_max + 1, nullptr);
  sgraph_.assign(2 * (nrhs_max + 1), nullptr);
  dgraph_.assign(nrhs_max + 1, nullptr);
  cgraph_.assign(2 * (nrhs_max + 1), nullptr);
}

FemContext::FemContext(const std::vector<FemContext*>& members, int nrhs_max, hipStream_t stream)
    : s_(stream), nrhs_max_(nrhs_max) {
  const int nop = (int)members.size();
  ANNEAL_CHECK(nop >= 1 && nop <= 64, "a group holds 1..64 problems");
  ANNEAL_CHECK(nrhs_max >= nop && nrhs_max <= 64, "group nrhs_max must be in [group size, 64]");
  for (FemContext* m : members) {
    ANNEAL_CHECK(m != nullptr, "null member context");
    ANNEAL_CHECK(!m->is_group(), "a group member must be a single problem");
    ANNEAL_CHECK(m->s_ == stream, "group members must use the group's stream");
  }
  const FemContext& m0 = *members[0];
  for (FemContext* m : members) {
    ANNEAL_CHECK(m->lv_.size() == m0.lv_.size(), "group members need the same multigrid levels");
    for (size_t l = 0; l < m0.lv_.size(); ++l) {
      const GridDims &a = m->lv_[l].g, &b = m0.lv_[l].g;
      ANNEAL_CHECK(a.nx == b.nx && a.ny == b.ny && a.nz == b.nz, "group members need the same grid shape");
    }
    ANNEAL_CHECK(std::memcmp(m->ke_, m0.ke_, sizeof(ke_)) == 0, "group members need the same voxel size and Poisson ratio");
  }
  try {
    members_ = members;
    std::memcpy(ke_, m0.ke_, sizeof(ke_));
    k64_ = m0.k64_;
    cheb_lmax = m0.cheb_lmax;
    nu_pre = std::max(2, m0.nu_pre);
    nu_coarse = std::max(2, m0.nu_coarse);
    nu_deep = m0.nu_deep > 0 ? std::max(2, m0.nu_deep) : 0;
    deep_escalate_iters = m0.deep_escalate_iters;
    deep_escalate_degree = std::max(2, m0.deep_escalate_degree);
    cheb_ratio = m0.cheb_ratio;
    coarsen_power = m0.coarsen_power;
    omega = m0.omega;
    inner_tol_floor = m0.inner_tol_floor;
    inner_tol_factor = m0.inner_tol_factor;
    inner_max = m0.inner_max;
    use_graphs = m0.use_graphs;
    fuse_first_from = 0;  // the unfused first step has no per-problem operator path
    tuned_ = true;
    lv_.resize(m0.lv_.size());
    for (size_t l = 0; l < lv_.size(); ++l) {
      Level& L = lv_[l];
      const Level& M = m0.lv_[l];
      L.g = M.g;
      L.h = M.h;
      L.k = M.k;
      L.kdiag = M.kdiag;
      L.tile = M.tile;
      L.mask = M.mask;  // member 0's: kernels read the tables below
      L.Ee = M.Ee;
      L.dinv = M.dinv;
      const size_t vb = (size_t)nrhs_max * 3 * L.g.ldn() * sizeof(float);
      L.x = (float*)alloc(vb);
      L.x2 = (float*)alloc(vb);
      L.b = (float*)alloc(vb);
      L.r = (float*)alloc(vb);
      ANNEAL_HIP(hipMemset(L.x, 0, vb));
      ANNEAL_HIP(hipMemset(L.x2, 0, vb));
      ANNEAL_HIP(hipMemset(L.b, 0, vb));
      ANNEAL_HIP(hipMemset(L.r, 0, vb));
      std::vector<OpEntry> tab(nop);
      for (int p = 0; p < nop; ++p) {
        const FemContext& m = *members[p];
        tab[p] = OpEntry{m.lv_[l].Ee, m.Ee64_, m.lv_[l].mask, m.lv_[l].dinv, m.Ainv_};
      }
      OpEntry* d = (OpEntry*)alloc(nop * sizeof(OpEntry));
      ANNEAL_HIP(hipMemcpy(d, tab.data(), nop * sizeof(OpEntry), hipMemcpyHostToDevice));
      optab_.push_back(d);
    }
    Ainv_ = m0.Ainv_;
    Ee64_ = m0.Ee64_;
    alloc_solver(nrhs_max);
  } catch (...) {
    for (void* p : owned_) (void)hipFree(p);
    owned_.clear();
    throw;
  }
}

void FemContext::set_rhs_count(int nrhs) {
  if (!is_group()) return;
  const int nop = (int)members_.size();
  ANNEAL_CHECK(nrhs >= nop && nrhs % nop == 0,
               "a group solves the same number of right-hand sides for every problem (nrhs = k * group size)");
  rpo_ = nrhs / nop;
}

void FemContext::drop_graphs() {
  ++graph_gen_;  // graphs recorded elsewhere from this context's kernels are stale too
  for (auto* set : {&vgraph_, &sgraph_, &dgraph_, &cgraph_})
    for (auto& g : *set)
      if (g) {
        (void)hipGraphExecDestroy(g);
        g = nullptr;
      }
  for (auto& kv : vgraph_bx_)
    if (kv.second) (void)hipGraphExecDestroy(kv.second);
  vgraph_bx_.clear();
}

template <typename F>
hipGraphExec_t FemContext::capture(F&& issue) {
  hipStream_t saved = s_;
  s_ = cap_s_;
  hipGraph_t g = nullptr;
  try {
    ANNEAL_HIP(hipStreamBeginCapture(cap_s_, hipStreamCaptureModeThreadLocal));
    issue();
    ANNEAL_HIP(hipStreamEndCapture(cap_s_, &g));
  } catch (...) {
    s_ = saved;
    hipGraph_t dummy = nullptr;
    (void)hipStreamEndCapture(cap_s_, &dummy);
    if (dummy) (void)hipGraphDestroy(dummy);
    throw;
  }
  s_ = saved;
  hipGraphExec_t exec = nullptr;
  const hipError_t e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  ANNEAL_HIP(e);
  return exec;
}

void FemContext::vcycle_graph(const float* b, float* x, int nrhs) {
  if (!use_graphs) {
    vcycle(0, b, x, nrhs);
    return;
  }
  const auto key = std::make_tuple(nrhs, (const void*)b, (void*)x);
  auto it = vgraph_bx_.find(key);
  if (it == vgraph_bx_.end()) it = vgraph_bx_.emplace(key, capture([&] { vcycle(0, b, x, nrhs); })).first;
  ANNEAL_HIP(hipGraphLaunch(it->second, s_));
}

void FemContext::set_learned(UNet* net, int level, double e_ref, int mode) {
  ANNEAL_CHECK(mode >= 0 && mode <= 2, "learned mode must be 0, 1 or 2");
  ANNEAL_CHECK(!is_group() || !net || mode == 0, "a group solves with the plain multigrid preconditioner");
  if (!net || mode == 0) {
    net_ = nullptr;
    net_mode_ = 0;
    return;
  }
  ANNEAL_CHECK(level >= 1 && level < (int)lv_.size(), "the network level must be a coarse multigrid level");
  ANNEAL_CHECK(e_ref > 0.0 && std::isfinite(e_ref), "e_ref must be positive");
  ANNEAL_CHECK(net->cin() == kNetCin && net->cout() == kNetCout, "the network must map 8 input to 3 output channels");
  const GridDims& g = lv_[level].g;
  if (!net_in_ || level != net_level_) {
    net_in_ = (uint16_t*)alloc((size_t)nrhs_max_ * g.nnode() * kNetCin * sizeof(uint16_t));
    net_out_ = (float*)alloc((size_t)nrhs_max_ * g.nnode() * kNetCout * sizeof(float));
  }
  if (!net_nrm_) net_nrm_ = (double*)alloc(4 * 64 * sizeof(double));
  if (!pz1_) {
    const size_t v32 = (size_t)nrhs_max_ * 3 * lv_[0].g.ldn() * sizeof(float);
    pz1_ = (float*)alloc(v32);
    pt_ = (float*)alloc(v32);
  }
  net_ = net;
  net_level_ = level;
  net_eref_ = e_ref;
  net_mode_ = mode;
}

void FemContext::learned_stage(const float* r, float* z, int nrhs) {
  ANNEAL_CHECK(net_, "no network attached");
  const float* src = r;
  for (int l = 1; l <= net_level_; ++l) {
    launch_restrict(src, lv_[l].b, lv_[l].mask, lv_[l - 1].g, lv_[l].g, nrhs, s_);
    src = lv_[l].b;
  }
  Level& N = lv_[net_level_];
  dot_partial(N.b, N.b, N.g.nnode(), N.g.ldn(), nrhs, part_, s_);
  finalize(part_, kDotBlocks, nrhs, net_nrm_, s_);
  net_pack(N.Ee, N.mask, N.b, net_nrm_, N.g, nrhs, net_eref_, net_in_, s_);
  net_->forward(net_in_, NetDims{nrhs, N.g.nx + 1, N.g.ny + 1, N.g.nz + 1}, net_out_);
  net_unpack(net_out_, N.mask, net_nrm_, N.g, nrhs, net_eref_, N.x, s_);
  for (int l = net_level_; l >= 1; --l) {
    float* dst = l == 1 ? z : lv_[l - 1].x;
    fill_zero(dst, (int64_t)nrhs * 3 * lv_[l - 1].g.ldn(), s_);
    launch_prolong_add(lv_[l].x, dst, lv_[l - 1].mask, lv_[l - 1].g, lv_[l].g, nrhs, s_);
  }
}

void FemContext::learned_pair(const double* F, const double* U, int nrhs, int level, double e_ref, uint16_t* in,
                              float* tgt) {
  ANNEAL_CHECK(!is_group(), "training pairs come from single problems");
  ANNEAL_CHECK(nrhs >= 1 && nrhs <= nrhs_max_, "nrhs out of range for this context");
  ANNEAL_CHECK(level >= 1 && level < (int)lv_.size(), "the network level must be a coarse multigrid level");
  ANNEAL_CHECK(e_ref > 0.0, "e_ref must be positive");
  if (!net_nrm_) net_nrm_ = (double*)alloc(4 * 64 * sizeof(double));
  const GridDims& g0 = lv_[0].g;
  for (int r = 0; r < nrhs; ++r)
    for (int c = 0; c < 3; ++c) cast_f64_f32(F + (size_t)(3 * r + c) * g0.ldn(), pr_ + (size_t)(3 * r + c) * g0.ldn(),
                                             g0.nnode(), s_);
  const float* src = pr_;
  for (int l = 1; l <= level; ++l) {
    launch_restrict(src, lv_[l].b, lv_[l].mask, lv_[l - 1].g, lv_[l].g, nrhs, s_);
    src = lv_[l].b;
  }
  Level& N = lv_[level];
  dot_partial(N.b, N.b, N.g.nnode(), N.g.ldn(), nrhs, part_, s_);
  finalize(part_, kDotBlocks, nrhs, net_nrm_, s_);
  net_pack(N.Ee, N.mask, N.b, net_nrm_, N.g, nrhs, e_ref, in, s_);
  net_target(U, g0, level, N.g, net_nrm_, nrhs, e_ref, tgt, s_);
}



void FemContext::vcycle_pcg(int nrhs) {
  auto issue = [&] { vcycle(0, pr_, pz_, nrhs, part_, st_); };
  if (!use_graphs) {
    issue();
    return;
  }
  if (!vgraph_[nrhs]) vgraph_[nrhs] = capture(issue);
  ANNEAL_HIP(hipGraphLaunch(vgraph_[nrhs], s_));
}

FemContext::~FemContext() {
  (void)hipStreamSynchronize(s_);
  drop_graphs();
  if (cap_s_) (void)hipStreamDestroy(cap_s_);
  if (ev_) (void)hipEventDestroy(ev_);
  if (ev2_) (void)hipEventDestroy(ev2_);
  for (void* p : owned_) (void)hipFree(p);
  if (host_red_) (void)hipHostFree(host_red_);
  if (host_up_) (void)hipHostFree(host_up_);
}

template <>
ApplyArgs<float> FemContext::args<float>(int l) const {
  const Level& L = lv_[l];
  ApplyArgs<float> a{};
  a.variant = L.tile;
  a.Ee = L.Ee;
  a.mask = L.mask;
  a.dinv = L.dinv;
  a.k = L.k;
  a.omega = omega;
  a.nx = L.g.nx;
  a.ny = L.g.ny;
  a.nz = L.g.nz;
  a.ldn = L.g.ldn();
  a.ops = optab(l);
  a.rpo = rpo_;
  return a;
}
template <>
ApplyArgs<double> FemContext::args<double>(int l) const {
  const Level& L = lv_[0];
  ApplyArgs<double> a{};
  a.variant = 0;
  a.Ee = Ee64_;
  a.mask = L.mask;
  a.dinv = L.dinv;
  a.k = k64_;
  a.omega = omega;
  a.nx = L.g.nx;
  a.ny = L.g.ny;
  a.nz = L.g.nz;
  a.ldn = L.g.ldn();
  a.ops = optab(0);
  a.rpo = rpo_;
  (void)l;
  return a;
}

void FemContext::issue_step(int nrhs, int slot, bool fused) {
  const GridDims& g0 = lv_[0].g;
  const int64_t n = g0.nnode(), ld = g0.ldn();
  ApplyArgs<float> af = args<float>(0);
  af.y = pq_;
  af.dot_part = part_;
  af.skip = st_;
  if (fused) {
    // p = z + beta p_old staged by the apply itself (the apply is bound by its
    // arithmetic, so the extra array costs less than a separate vector pass)
    af.u = pdir(slot ^ 1);
    af.b = pz_;
    af.pout = pdir(slot);
    launch_apply(af, kApplyDir, nrhs, s_);
  } else {
    af.u = pdir(slot);
    launch_apply(af, kApply, nrhs, s_);
  }
  finalize_st(part_, apply_blocks(g0.nx, g0.ny, g0.nz, lv_[0].tile), nrhs, st_, kPQ, 1, s_);  // pq, alpha
  // |r|^2 to the host; a rhs that meets its inner tolerance is marked done on the device
  pcg_update_converge(px_, pr_, pdir(slot), pq_, st_, n, ld, nrhs, part_, s_, host_red_dev_ + rr_slot(slot),
                      host_red_dev_ + done_slot(slot));
}

void FemContext::issue_direction(int nrhs) {
  const GridDims& g0 = lv_[0].g;
  // the cycle's last fine-level smoothing kernel leaves the <r, z> partials on the apply grid
  vcycle(0, pr_, pz_, nrhs, part_, st_);
  finalize_st(part_, apply_blocks(g0.nx, g0.ny, g0.nz, lv_[0].tile), nrhs, st_, kRZnew, 2, s_);  // rz, beta
}

void FemContext::pcg_step(int nrhs, int slot) {
  if (!use_graphs) {
    issue_step(nrhs, slot, false);
    return;
  }
  hipGraphExec_t& g = sgraph_[2 * nrhs + slot];
  if (!g) g = capture([&] { issue_step(nrhs, slot, false); });
  ANNEAL_HIP(hipGraphLaunch(g, s_));
}

void FemContext::pcg_direction(int nrhs) {
  if (!use_graphs) {
    issue_direction(nrhs);
    return;
  }
  if (!dgraph_[nrhs]) dgraph_[nrhs] = capture([&] { issue_direction(nrhs); });
  ANNEAL_HIP(hipGraphLaunch(dgraph_[nrhs], s_));
}

void FemContext::pcg_direction_step(int nrhs, int slot) {
  if (!use_graphs) {
    issue_direction(nrhs);
    issue_step(nrhs, slot, true);
    return;
  }
  hipGraphExec_t& g = cgraph_[2 * nrhs + slot];
  if (!g) g = capture([&] {
      issue_direction(nrhs);
      issue_step(nrhs, slot, true);
    });
  ANNEAL_HIP(hipGraphLaunch(g, s_));
}

void FemContext::set_Ee(const double* Ee_dev) {
  ANNEAL_CHECK(!is_group(), "set the stiffness on the problems of a group, not on the group");
  const GridDims& g0 = lv_[0].g;
  copy_cast_f64(Ee_dev, Ee64_, lv_[0].Ee, g0.nelem(), s_);
  for (size_t l = 1; l < lv_.size(); ++l)
    launch_coarsen_Ee(lv_[l - 1].Ee, lv_[l].Ee, lv_[l - 1].g, lv_[l].g, (float)coarsen_power, s_);
  DiagLevels d{};
  d.nlev = (int)lv_.size();
  ANNEAL_CHECK(d.nlev <= DiagLevels::kMax, "too many multigrid levels");
  for (int l = 0; l < d.nlev; ++l) {
    const Level& L = lv_[l];
    d.Ee[l] = L.Ee;
    d.mask[l] = L.mask;
    d.dinv[l] = L.dinv;
    d.nx[l] = L.g.nx;
    d.ny[l] = L.g.ny;
    d.nz[l] = L.g.nz;
    d.kdiag[l] = L.kdiag;
  }
  launch_inv_diag(d, s_);
  coarse_stale_ = true;
}

void FemContext::rebuild_coarse() {
  if (is_group()) {
    // each member keeps its own rebuild policy; coarse_valid_ stays false so every solve looks again
    for (FemContext* m : members_)
      if (!m->coarse_valid_ || (m->coarse_stale_ && m->rebuild_each_solve)) m->rebuild_coarse();
    return;
  }
  const Level& C = lv_.back();
  dense_assemble_invert(C.Ee, C.mask, ke_coarse_dev_, Ainv_, Awork_, C.g.nx, C.g.ny, C.g.nz, s_);
  coarse_stale_ = false;
  coarse_valid_ = true;
}

void FemContext::apply_f32(const float* u, float* y, int nrhs) {
  set_rhs_count(nrhs);
  ApplyArgs<float> a = args<float>(0);
  a.u = u;
  a.y = y;
  launch_apply(a, kApply, nrhs, s_);
}
void FemContext::smooth_step_f32(const float* u, const float* b, float* y, int nrhs) {
  set_rhs_count(nrhs);
  const double lmax = cheb_lmax, lmin = lmax / cheb_ratio;
  const double th = 0.5 * (lmax + lmin), de = 0.5 * (lmax - lmin), sg = th / de;
  const double rn = 1.0 / (2.0 * sg - 1.0 / sg);  // the recurrence's k = 1 coefficients (as in vcycle)
  ApplyArgs<float> a = args<float>(0);
  a.u = u;
  a.b = b;
  a.prev = y;
  a.y = y;
  a.c1 = float(rn / sg);
  a.c2 = float(2.0 * rn / de);
  launch_apply(a, kCheb, nrhs, s_);
}
void FemContext::apply_f64(const double* u, double* y, int nrhs) {
  set_rhs_count(nrhs);
  ApplyArgs<double> a = args<double>(0);
  a.u = u;
  a.y = y;
  launch_apply(a, kApply, nrhs, s_);
}

bool FemContext::maybe_escalate(const int* its, int nrhs) {
  if (deep_escalate_iters <= 0 || nu_deep != 0 || lv_.size() < 4 || nrhs < 1) return false;
  long sum = 0;
  for (int r = 0; r < nrhs; ++r) sum += its[r];
  if (sum <= (long)deep_escalate_iters * nrhs) return false;
  nu_deep = deep_escalate_degree;
  drop_graphs();  // the smoothing degree is baked into captured V-cycles
  return true;
}

void FemContext::set_level_tile(int level, int variant) {
  ANNEAL_CHECK(!is_group(), "a group uses the tile shapes of its first problem");
  ANNEAL_CHECK(level >= 0 && level < (int)lv_.size(), "level out of range");
  ANNEAL_CHECK(variant >= 0 && variant < apply_variants(), "unknown apply tile variant");
  drop_graphs();
  lv_[level].tile = variant;
  tuned_ = true;  // a pinned shape is kept: no sweep when the first operator is set
}

// Tile shapes chosen by the timing sweep, per device and multigrid hierarchy
// (the level dimensions), shared by every context of the process: the
// apply's time does not depend on the stiffness values, so a context on a
// grid that was tuned before (the next pass of a volume-fraction search, the
// next run of a benchmark) takes the shapes without a sweep.
namespace {
std::mutex g_tiles_mu;
std::map<std::vector<int>, std::vector<int>> g_tiles;
}  // namespace

std::vector<int> FemContext::tile_key() const {
  int dev = 0;
  ANNEAL_HIP(hipGetDevice(&dev));
  std::vector<int> key{dev};
  for (const auto& L : lv_) {
    key.push_back(L.g.nx);
    key.push_back(L.g.ny);
    key.push_back(L.g.nz);
  }
  return key;
}

void FemContext::pick_tiles() {
  const std::vector<int> key = tile_key();
  {
    std::lock_guard<std::mutex> lk(g_tiles_mu);
    auto it = g_tiles.find(key);
    if (it != g_tiles.end()) {
      for (size_t l = 0; l < lv_.size(); ++l) lv_[l].tile = it->second[l];
      tuned_ = true;
      return;
    }
  }
  autotune(-1, nullptr);  // stores its choice
}

int FemContext::autotune(int variant, float* us_out) {
  ANNEAL_CHECK(!is_group(), "a group uses the tile shapes of its first problem");
  const int nv = apply_variants();
  drop_graphs();  // tile shapes are baked into captured V-cycles
  if (variant >= 0) {
    ANNEAL_CHECK(variant < nv, "unknown apply tile variant");
    for (auto& L : lv_) L.tile = variant;
    return variant;
  }
  // time the FP32 apply (one rhs) of every tile shape on every level with a
  // stencil; keep the fastest per level
  hipEvent_t e0, e1;
  ANNEAL_HIP(hipEventCreate(&e0));
  ANNEAL_HIP(hipEventCreate(&e1));
  for (size_t l = 0; l + 1 < lv_.size() || (l == 0 && lv_.size() == 1); ++l) {
    Level& L = lv_[l];
    float best = 1e30f;
    int arg = 0;
    for (int v = 0; v < nv; ++v) {
      L.tile = v;
      ApplyArgs<float> a = args<float>((int)l);
      a.u = l == 0 ? pp_ : L.x;
      a.y = l == 0 ? pq_ : L.r;
      launch_apply(a, kApply, 1, s_);  // warm
      float us = 1e30f;
      for (int round = 0; round < 3; ++round) {  // best of three: small grids are noisy
        ANNEAL_HIP(hipEventRecord(e0, s_));
        const int reps = 8;
        for (int i = 0; i < reps; ++i) launch_apply(a, kApply, 1, s_);
        ANNEAL_HIP(hipEventRecord(e1, s_));
        ANNEAL_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        ANNEAL_HIP(hipEventElapsedTime(&ms, e0, e1));
        us = std::min(us, 1e3f * ms / reps);
      }
      if (l == 0 && us_out) us_out[v] = us;
      if (us < best) {
        best = us;
        arg = v;
      }
    }
    L.tile = arg;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  tuned_ = true;
  std::vector<int> tiles;
  for (const auto& L : lv_) tiles.push_back(L.tile);
  {
    std::lock_guard<std::mutex> lk(g_tiles_mu);
    g_tiles[tile_key()] = tiles;
  }
  return lv_[0].tile;
}

void FemContext::vcycle(int l, const float* b, float* x, int nrhs, double* dot_part, const double* skip) {
  Level& L = lv_[l];
  if (l == (int)lv_.size() - 1) {
    dense_solve(Ainv_, b, x, (int)L.g.nnode(), L.g.ldn(), nrhs, s_, optab(l), rpo_, skip);
    return;
  }
  const int64_t n = L.g.nnode(), ld = L.g.ldn();
  const int nu = l == 0 ? nu_pre : (l >= 2 && nu_deep > 0 ? nu_deep : nu_coarse);  // smoothing degree here
  // Chebyshev recurrence (three-term form): x_{k+1} = x_k + c1_k (x_k - x_{k-1}) + c2_k D^-1 (b - K x_k)
  const double lmax = cheb_lmax, lmin = lmax / cheb_ratio;
  const double th = 0.5 * (lmax + lmin), de = 0.5 * (lmax - lmin), sg = th / de;
  double c1[16], c2[16];
  double rho = 1.0 / sg;
  for (int k = 1; k < nu && k < 16; ++k) {
    const double rn = 1.0 / (2.0 * sg - rho);
    c1[k] = rn * rho;
    c2[k] = 2.0 * rn / de;
    rho = rn;
  }
  // 2*nu - 1 buffer swaps in total: start in x2 so the result lands in x.
  float* cur = L.x2;
  float* nxt = x;
  ApplyArgs<float> a = args<float>(l);
  a.b = b;
  a.skip = skip;
  if (nu >= 2 && l >= fuse_first_from) {
    // steps 1 and 2 fused: x1 = omega D^-1 b is formed inside the kernel,
    // x2 lands where the unfused sequence would put it; step 3 forms its
    // previous iterate x1 at the node as well
    a.u = nullptr;
    a.prev = nullptr;
    a.y = nxt;
    a.omega = float(1.0 / th);
    a.c1 = float(c1[1]);
    a.c2 = float(c2[1]);
    launch_apply(a, kCheb0, nrhs, s_);
    std::swap(cur, nxt);
    for (int k = 2; k < nu; ++k) {
      a.u = cur;
      a.prev = (k == 2) ? nullptr : nxt;
      a.prev_j0 = (k == 2) ? 1 : 0;
      a.omega0 = float(1.0 / th);
      a.y = nxt;
      a.c1 = float(c1[k]);
      a.c2 = float(c2[k]);
      launch_apply(a, kCheb, nrhs, s_);
      std::swap(cur, nxt);
    }
    a.prev_j0 = 0;
  } else {
    ANNEAL_CHECK(!is_group(), "a group smooths with the fused first step (degree >= 2)");
    jacobi0(b, cur, L.dinv, L.mask, float(1.0 / th), n, ld, nrhs, s_, skip);
    for (int k = 1; k < nu; ++k) {
      a.u = cur;
      a.prev = (k == 1) ? nullptr : nxt;
      a.y = nxt;
      a.c1 = float(c1[k]);
      a.c2 = float(c2[k]);
      launch_apply(a, kCheb, nrhs, s_);
      std::swap(cur, nxt);
    }
  }
  a.u = cur;
  a.y = L.r;
  launch_apply(a, kResid, nrhs, s_);
  Level& C = lv_[l + 1];
  launch_restrict(L.r, C.b, C.mask, L.g, C.g, nrhs, s_, optab(l + 1), rpo_, skip);
  vcycle(l + 1, C.b, C.x, nrhs, nullptr, skip);
  launch_prolong_add(C.x, cur, L.mask, L.g, C.g, nrhs, s_, optab(l), rpo_, skip);
  a.u = cur;
  a.y = nxt;
  a.omega = float(1.0 / th);
  a.dot_part = (l == 0 && nu == 1) ? dot_part : nullptr;  // the last kernel of the cycle forms <b, x>
  launch_apply(a, kJacobi, nrhs, s_);
  std::swap(cur, nxt);
  for (int k = 1; k < nu; ++k) {
    a.u = cur;
    a.prev = nxt;
    a.y = nxt;
    a.c1 = float(c1[k]);
    a.c2 = float(c2[k]);
    a.dot_part = (l == 0 && k == nu - 1) ? dot_part : nullptr;
    launch_apply(a, kCheb, nrhs, s_);
    std::swap(cur, nxt);
  }
  if (cur != x) ANNEAL_HIP(hipMemcpyAsync(x, cur, (size_t)nrhs * 3 * ld * sizeof(float), hipMemcpyDeviceToDevice, s_));
}

int FemContext::solve(const double* F, double* U, int nrhs, double rtol, int maxiter, bool use_x0, int* iters_out) {
  ANNEAL_CHECK(nrhs >= 1 && nrhs <= nrhs_max_, "nrhs out of range for this context");
  set_rhs_count(nrhs);
  if (!coarse_valid_ || (coarse_stale_ && rebuild_each_solve)) rebuild_coarse();
  const GridDims& g0 = lv_[0].g;
  const int64_t n = g0.nnode(), ld = g0.ldn();
  const size_t v64 = (size_t)nrhs * 3 * ld * sizeof(double);
  if (!use_x0) ANNEAL_HIP(hipMemsetAsync(U, 0, v64, s_));
  // |F| is read after the first residual's wait: one host round trip for both
  dot(F, F, n, ld, nrhs, part_, red_ + 128, s_, host_red_dev_ + 128);
  std::vector<double> fnorm(nrhs), rnorm(nrhs), tol2(nrhs), sc(nrhs);
  std::vector<int> its(nrhs, 0);
  std::vector<char> done(nrhs, 0), idone(nrhs, 0);
  std::vector<double> hst((size_t)nrhs * kSt, 0.0);
  const int nblk_apply = apply_blocks(g0.nx, g0.ny, g0.nz, lv_[0].tile);
  int status = 1;
  int total = 0;
  bool bad = false;
  for (int outer = 0; outer < 50 && !bad; ++outer) {
    // true residual in FP64
    ApplyArgs<double> a = args<double>(0);
    a.u = U;
    a.b = F;
    a.y = r64_;
    a.dot_part = part_;  // |r|^2 partials from the residual kernel itself
    launch_apply(a, kResid, nrhs, s_);
    finalize(part_, apply_blocks(g0.nx, g0.ny, g0.nz, kApplyVariant64), nrhs, red_, s_, host_red_dev_);
    // The host decisions of a pass, from |r|^2 and |F|^2 in host_red_ (after a wait).
    bool all = true;
    auto decide = [&]() {
      if (outer == 0)
        for (int r = 0; r < nrhs; ++r) fnorm[r] = std::sqrt(host_red_[128 + r]);
      all = true;
      for (int r = 0; r < nrhs; ++r) {
        rnorm[r] = std::sqrt(host_red_[r]);
        if (!std::isfinite(rnorm[r])) bad = true;
        done[r] = (rnorm[r] <= rtol * fnorm[r]) ? 1 : 0;
        all = all && done[r];
        const double want = done[r] ? 1.0 : inner_tol_factor * rtol * fnorm[r] / rnorm[r];
        const double t = std::max(want, inner_tol_floor);
        tol2[r] = t * t;
      }
    };
    // The first pass (a fresh start almost never meets the tolerance) is set up
    // on the device from the same norms and the host decides after its first
    // PCG step's wait; a pass after that usually ends the solve, so the host
    // waits for its residual first.
    const bool optimistic = outer == 0;
    bool pending = optimistic;
    if (optimistic) {
      pass_setup(red_, red_ + 128, rtol, inner_tol_factor, inner_tol_floor, scale_, st_, nrhs, s_);
      for (int r = 0; r < nrhs; ++r) idone[r] = 0;
    } else {
      ANNEAL_HIP(hipStreamSynchronize(s_));
      decide();
      if (bad) break;
      if (all) {
        status = 0;
        break;
      }
      if (total >= maxiter) break;
      // inner FP32 PCG on the residual scaled to unit norm
      for (int r = 0; r < nrhs; ++r) {
        sc[r] = done[r] ? 0.0 : 1.0 / rnorm[r];
        for (int k = 0; k < kSt; ++k) hst[r * kSt + k] = 0.0;
        hst[r * kSt + kDone] = done[r] ? 1.0 : 0.0;
        hst[r * kSt + kTol2] = tol2[r];
        idone[r] = done[r];
      }
      std::copy(sc.begin(), sc.end(), host_up_);
      std::copy(hst.begin(), hst.end(), host_up_ + 64);
      upload(host_up_dev_, scale_, nrhs, host_up_dev_ + 64, st_, (int)hst.size(), s_);
    }
    // the pass's decisions once its residual norms have reached the host
    auto resolve = [&]() {
      pending = false;
      decide();
      for (int r = 0; r < nrhs; ++r) {
        idone[r] = done[r];
        hst[r * kSt + kDone] = done[r] ? 1.0 : 0.0;
      }
    };
    to_f32_scaled(r64_, pr_, scale_, n, ld, nrhs, s_);
    ANNEAL_HIP(hipMemsetAsync(px_, 0, (size_t)nrhs * 3 * ld * sizeof(float), s_));
    if (net_ && net_mode_ == 1) learned_warm_start(nrhs);
    if (!precondition(nrhs)) {
      dot_partial(pr_, pz_, n, ld, nrhs, part_, s_);
      finalize_st(part_, kDotBlocks, nrhs, st_, kRZ, 0, s_);
    } else {
      finalize_st(part_, nblk_apply, nrhs, st_, kRZ, 0, s_);
    }
    pcg_dir(pp_, pz_, st_, n, ld, nrhs, 1, s_);
    // rr of the last two iterations (per rhs) to predict convergence: when the
    // next residual is not expected to meet the tolerance, the next V-cycle,
    // dot and direction update (they read r and write z, p, rz, beta only) are
    // queued before the host waits for rr, so the device never idles on it.
    // The pass starts from rr = 1 (the residual scaled to unit norm); before two
    // rr are known the prediction takes the first iteration's contraction to be
    // kFastFirst, about the fastest seen on the cantilever schedule (0.003-0.47,
    // median 0.009 over 120 passes), and the second one's to equal the first's,
    // which is faster than the median second (0.04): the queued work is
    // rarely wasted.
    constexpr double kFastFirst = 0.003;
    std::vector<double> rr1(nrhs, 1.0), rr2(nrhs, 1.0 / kFastFirst);
    // The plain multigrid iteration runs as one graph: the direction of step k
    // followed by step k + 1 (one launch per iteration).  A rhs whose step meets
    // its tolerance is marked done on the device, so a step queued before the
    // host has seen that leaves it unchanged; the queued step stores its |r|^2
    // and done flags in the other host slot, not over the ones the host reads.
    const bool flexible = net_ && net_mode_ == 2;
    const bool merged = !flexible && lv_.size() > 1;
    bool step_queued = false;  // step it was queued with the previous direction
    int slot = 0;              // host slot of the step the host waits for next
    for (int it = 0; it < inner_max && total < maxiter; ++it) {
      if (!step_queued) {
        pcg_step(nrhs, slot);
        ANNEAL_HIP(hipEventRecord(ev_, s_));
      }
      step_queued = false;
      // the first step of an optimistic pass is followed by another unless the
      // start already met the tolerance (the first pass took 2-14 steps on every
      // solve of the cantilever schedule): queue it before the tolerances are known
      bool speculate = true;
      for (int r = 0; r < nrhs && speculate && !pending; ++r) {
        if (idone[r]) continue;
        const double rate = rr1[r] / rr2[r];
        if (rr1[r] * rate <= 2.0 * tol2[r]) speculate = false;  // may converge now: do not queue work
      }
      // total_after: the step count once this step is counted
      auto next_direction = [&](int total_after) {
        if (merged) {
          // with the next step when the loop would run it
          if (it + 1 < inner_max && total_after < maxiter) {
            pcg_direction_step(nrhs, slot ^ 1);
            ANNEAL_HIP(hipEventRecord(ev2_, s_));
            step_queued = true;
          } else {
            pcg_direction(nrhs);
          }
          return;
        }
        const bool fused = precondition(nrhs);
        if (!fused) dot_partial(pr_, pz_, n, ld, nrhs, part_, s_, st_);
        const int nblk_rz = fused ? nblk_apply : kDotBlocks;
        if (flexible) {
          finalize_st(part_, nblk_rz, nrhs, st_, kRZnew, 4, s_);  // rz for the next alpha
          dot_partial(pz_, pq_, n, ld, nrhs, part_, s_);
          finalize_st(part_, kDotBlocks, nrhs, st_, kZQ, 3, s_);  // beta = -<z, K p> / <p, K p>
        } else {
          finalize_st(part_, nblk_rz, nrhs, st_, kRZnew, 2, s_);  // rz_new, beta
        }
        pcg_dir(pp_, pz_, st_, n, ld, nrhs, 0, s_);
      };
      if (speculate) next_direction(total + 1);
      ANNEAL_HIP(hipEventSynchronize(ev_));
      if (pending) {
        resolve();
        if (bad || all) break;  // the device skipped every rhs already within the tolerance
      }
      ++total;
      bool iall = true;
      for (int r = 0; r < nrhs; ++r) {
        if (idone[r]) continue;
        ++its[r];
        const double rr = host_red_[rr_slot(slot) + r];
        rr2[r] = rr1[r];
        rr1[r] = rr;
        if (!std::isfinite(rr)) bad = true;
        if (host_red_[done_slot(slot) + r] != 0.0) idone[r] = 1;  // the device's test |r|^2 <= tol2
        iall = iall && idone[r];
      }
      if (bad || iall) break;
      if (!speculate) next_direction(total);
      if (step_queued) {
        std::swap(ev_, ev2_);
        slot ^= 1;
      }
    }
    if (pending) {  // no PCG step ran (maxiter reached): decide from the residual alone
      ANNEAL_HIP(hipStreamSynchronize(s_));
      resolve();
    }
    if (bad) break;
    if (all) {
      status = 0;
      break;
    }
    for (int r = 0; r < nrhs; ++r) host_up_[r] = done[r] ? 0.0 : rnorm[r];
    upload(host_up_dev_, scale_, nrhs, nullptr, nullptr, 0, s_);
    axpy_f32_to_f64(px_, U, scale_, n, ld, nrhs, s_);
  }
  if (bad) status = 2;
  ANNEAL_HIP(hipStreamSynchronize(s_));
  for (int r = 0; r < nrhs; ++r) iters_out[r] = its[r];
  maybe_escalate(its.data(), nrhs);
  return status;
}

// z1 = gamma P N(R r) with gamma = <r, z1> / <z1, K z1> (the energy-optimal
// length of the network step), t = r - K z1.  Reads r, writes z1 and t.
void FemContext::scaled_network_step(const float* r, float* z1, float* t, int nrhs) {
  const GridDims& g0 = lv_[0].g;
  const int64_t n = g0.nnode(), ld = g0.ldn();
  learned_stage(r, z1, nrhs);
  ApplyArgs<float> a = args<float>(0);
  a.u = z1;
  a.y = pt_;
  a.dot_part = part_;
  launch_apply(a, kApply, nrhs, s_);  // pt = K z1, partials of <z1, K z1>
  const int nblk = apply_blocks(g0.nx, g0.ny, g0.nz, lv_[0].tile);
  finalize(part_, nblk, nrhs, net_nrm_ + 64, s_);
  dot_partial(r, z1, n, ld, nrhs, part_, s_);
  finalize(part_, kDotBlocks, nrhs, net_nrm_ + 128, s_);
  net_gamma(net_nrm_ + 128, net_nrm_ + 64, nrhs, net_nrm_ + 192, s_);
  net_scaled_resid(r, pt_, z1, t, net_nrm_ + 192, ld, nrhs, s_);  // t = r - gamma K z1, z1 *= gamma
}

bool FemContext::precondition(int nrhs) {
  if (!net_ || net_mode_ != 2) {
    vcycle_pcg(nrhs);
    // the cycle's last smoothing kernel leaves the <r, z> partials in part_;
    // a one-level hierarchy is a dense solve only and leaves none
    return lv_.size() > 1;
  }
  // multiplicative two-stage preconditioner: z = z1 + V(r - K z1)
  const int64_t nv = (int64_t)nrhs * 3 * lv_[0].g.ldn();
  scaled_network_step(pr_, pz1_, pt_, nrhs);
  vcycle_graph(pt_, pz_, nrhs);
  add_f32(pz_, pz1_, nv, s_);
  return false;
}

void FemContext::learned_warm_start(int nrhs) {
  // x = z1, r = r - K z1 before the (plain) inner PCG starts
  const size_t bytes = (size_t)nrhs * 3 * lv_[0].g.ldn() * sizeof(float);
  scaled_network_step(pr_, pz1_, pr_, nrhs);
  ANNEAL_HIP(hipMemcpyAsync(px_, pz1_, bytes, hipMemcpyDeviceToDevice, s_));
}

}  // namespace anneal
