// 3-D convolutional U-Net on node grids: BF16 MFMA implicit-GEMM convolutions
// (forward, data gradient, weight gradient) with FP32 accumulation, used as a
// learned coarse-grid stage of the flexible PCG.
//
// Tensors are channels-last: voxel v = i + nx (j + ny k) (+ sample * nx ny nz),
// element [v][c] at v * pitch + c, bf16.  A "view" is (pointer, channels,
// pitch), so a convolution can read or write a channel slice of a wider
// buffer (the decoder's concatenations are slices, not copies).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "common.hpp"

namespace anneal {

using bf16 = uint16_t;  // storage type on the host side; kernels use __bf16

struct NetDims {
  int B = 0, nx = 0, ny = 0, nz = 0;
  int64_t vox() const { return (int64_t)B * nx * ny * nz; }
  bool operator==(const NetDims& o) const { return B == o.B && nx == o.nx && ny == o.ny && nz == o.nz; }
};
// Node-aligned coarsening: coarse node I sits on fine node 2I.
inline int coarse_n(int n) { return n / 2 + 1; }

struct View {
  bf16* p = nullptr;
  int c = 0, pitch = 0;
  View sub(int c0, int cn) const { return View{p + c0, cn, pitch}; }
};

// One convolution layer: 3x3x3 (taps 27, zero padding 1) or 1x1x1 (taps 1).
// Weights W[co][tap][ci] (co < cout real) are kept FP32 (master) in a padded
// [cout_pad][kp] layout, k = tap * cin + ci, with kp = round_up(taps * cin, 64);
// wt is their bf16 copy and wd the data-gradient kernel
// wd[ci][tap' * cout_pad + co] = W[co][26 - tap'][ci] (taps reversed).
struct ConvLayer {
  int taps = 27, cin = 0, cout = 0, cout_pad = 0, kp = 0, kdp = 0;
  bool relu = true;
  float *w = nullptr, *b = nullptr, *gw = nullptr, *gb = nullptr;  // FP32 master and gradients
  float *mw = nullptr, *vw = nullptr, *mb = nullptr, *vb = nullptr;  // Adam moments
  bf16 *wt = nullptr, *wd = nullptr;
  int64_t nparams() const { return (int64_t)cout * taps * cin + cout; }
};

struct ConvK;  // convolution launch arguments (unet.hip)

class UNet {
 public:
  // cin input channels (multiple of 8), cout output channels (<= 16),
  // widths[levels] feature channels per level (multiples of 16).
  UNet(int cin, int cout, const std::vector<int>& widths, hipStream_t s);
  ~UNet();
  UNet(const UNet&) = delete;
  UNet& operator=(const UNet&) = delete;

  int64_t nparams() const;
  // Flat parameter vector, layer by layer: W[co][tap][ci] then b[co].
  void get_params(float* host) const;
  void set_params(const float* host);
  void get_grads(float* host) const;  // same layout as get_params
  // out[v][cout] (FP32) = net(in[v][cin]) for B samples of nx x ny x nz nodes.
  void forward(const bf16* in, const NetDims& d, float* out);
  // Accumulates parameter gradients for dL/dout (FP32 [v][cout]) of the last
  // forward (same shape).  zero_grad() clears them.
  void backward(const float* dout);
  void zero_grad();
  void adam(float lr, float beta1, float beta2, float eps, int step, float weight_decay);
  // Sum of squared parameter gradients (device scalar, FP64) for clipping/diagnostics.
  void grad_sqnorm(double* out_dev);
  void scale_grad(const double* scale_dev);
  hipStream_t stream() const { return s_; }
  int cin() const { return cin_; }
  int cout() const { return cout_; }
  int levels() const { return (int)widths_.size(); }

 private:
  struct Plan {
    NetDims d[8];
    // forward activations per level
    std::vector<bf16*> owned;
    View in, a[8], cat[8], x[8], e[8], dec[8], skip[8], bott, bott_a;
    bf16* dout16 = nullptr;
    float* out = nullptr;
    // gradient buffers mirroring the activations
    View ga[8], gcat[8], gx[8], ge[8], gdec[8], gbott, gbott_a, gd0;
    float* wpart = nullptr;  // weight-gradient partials
    size_t wpart_elems = 0;
    int64_t nchunk_max = 0;
  };
  void make_plan(const NetDims& d0);
  void free_plan();
  void conv_fwd(ConvLayer& L, const View& x, const NetDims& d, const View& y, float* yf, int cf);
  void conv_dgrad(ConvLayer& L, const View& dy, const NetDims& d, const View& dx, const View* mask, bool accum);
  void conv_wgrad(ConvLayer& L, const View& x, const View& dy, const NetDims& d);
  void refresh_bf16(ConvLayer& L);
  void conv_dispatch(ConvK a, int n16);
  float* kpart_ = nullptr;  // split-K partial sums
  size_t kpart_elems_ = 0;

  hipStream_t s_;
  int cin_, cout_;
  std::vector<int> widths_;
  std::vector<ConvLayer> enc_a_, enc_b_, dec_a_, dec_b_;
  ConvLayer out_;
  std::vector<void*> owned_;
  Plan plan_;
  bool have_plan_ = false;
  bool fwd_valid_ = false;
  double* red_ = nullptr;
  double* part_ = nullptr;
  void* alloc(size_t bytes);
};

// ---- grid transfer and data preparation (also used by the solver) ----------
// y[I] = s * sum_d w(d) x[2I + d], w(d) = prod(1 - |d_i| / 2); out-of-range fine nodes skipped.
void net_restrict(const View& x, const NetDims& fine, const View& y, const NetDims& coarse, float s,
                  const View* mask, bool accum, hipStream_t st);
// y[i] (+)= s * sum over the coarse parents of i with trilinear weights (transpose of net_restrict with s = 1).
void net_prolong(const View& x, const NetDims& coarse, const View& y, const NetDims& fine, float s, const View* mask,
                 bool accum, hipStream_t st);

// Per-sample relative squared error: loss = mean_b sum (o - t)^2 / sum t^2 and
// its gradient dout.  o, t: FP32 [B][V][c].  loss_dev: one double.
void net_rel_mse(const float* o, const float* t, const NetDims& d, int c, float* dout, double* loss_dev, double* work,
                 hipStream_t st);

}  // namespace anneal
