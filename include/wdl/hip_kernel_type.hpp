// hip_kernel_type.hpp -- header-only KernelType objects backed by libwdl (C ABI).
//
// The reference's WassersteinRegression / Problem / losses take their
// convolution operator as a duck-typed template parameter (kernels.h,
// instantiated at main_dictionary_learning.cpp:168-186).  The classes here
// provide the same members, so
//     typedef wdl::HipGaussian2DKernel KernelType;               // kernels.h:630-708
//     typedef wdl::HipLogSignArrayGaussian2DKernel KernelType;   // kernels.h:1034-1202
//     typedef wdl::HipGaussian1DKernel KernelType;               // kernels.h:501-556
//     typedef wdl::HipGaussian3DKernel KernelType;               // kernels.h:712-815
// replaces Gaussian2DKernel / GaussianSSE2DKernel / GaussianHalide2DKernel or
// LogSignArrayGaussian2DKernel without touching the reference's code.
//
// Members used by the reference: N, gamma, (gamma, W, H) constructor, default
// constructor, copy/assign (inverseWasserstein.h:37, :48), convolve /
// convolveAdjoint (:568, :575, :611, :619, :655, :661), operator()(i, j)
// (:112), and in the log build log_convolve*, log_convolve_signArray*
// (:702-800).  Copies share one device object (reference counted), as the
// Halide kernels share their buffers on copy (kernels.h:1680-1702).
//
// The reference's operators return void and have no error channel; a failing
// call here prints wdl_last_error() and aborts.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstdlib>
#include <memory>

#include "../wdl_c.h"

namespace wdl {

namespace detail {

inline void check(int status, const char* what) {
  if (status != WDL_OK) {
    std::fprintf(stderr, "libwdl: %s failed (%d): %s\n", what, status, wdl_last_error());
    std::abort();
  }
}

// One context (device WDL_DEVICE or 0, one stream) shared by every kernel object.
inline std::shared_ptr<wdl_ctx> shared_context() {
  static std::shared_ptr<wdl_ctx> ctx = [] {
    const char* env = std::getenv("WDL_DEVICE");
    wdl_ctx* c = nullptr;
    check(wdl_ctx_create(env ? std::atoi(env) : 0, &c), "wdl_ctx_create");
    return std::shared_ptr<wdl_ctx>(c, [](wdl_ctx* p) { wdl_ctx_destroy(p); });
  }();
  return ctx;
}

}  // namespace detail

// Drop-in for Gaussian2DKernel (kernels.h:630-708) and its SSE/AVX and Halide
// variants: R = K_H . U . K_W per image, FP64 in and out, FP32 on the GPU.
class HipGaussian2DKernel {
 public:
  // kernels.h:632-641 (default regularization 25).
  explicit HipGaussian2DKernel(double regularization = 25., int W = 0, int H = 0)
      : gamma(regularization), W_(W), H_(H), N((size_t)W * H) {
    if (W > 0 && H > 0) {
      ctx_ = detail::shared_context();
      wdl_gauss2d* k = nullptr;
      detail::check(wdl_gauss2d_create(ctx_.get(), W, H, regularization, 0, &k),
                    "wdl_gauss2d_create");
      k_ = std::shared_ptr<wdl_gauss2d>(k, [](wdl_gauss2d* p) { wdl_gauss2d_destroy(p); });
    }
  }

  // kernels.h:656-662: entry (i, j) of the 2-D kernel.
  double operator()(int i, int j) const {
    const int x0 = i % (int)W_, y0 = i / (int)W_, x1 = j % (int)W_, y1 = j / (int)W_;
    const double d2 = (double)(x0 - x1) * (x0 - x1) + (double)(y0 - y1) * (y0 - y1);
    return std::max(1e-250, std::exp(-d2 / gamma));
  }

  // kernels.h:669-699
  void convolve(const double* u, double* result, int nvectors = 1) const {
    detail::check(wdl_gauss2d_convolve(k_.get(), u, result, nvectors), "wdl_gauss2d_convolve");
  }
  // kernels.h:700-702: the operator is symmetric.
  void convolveAdjoint(const double* u, double* result, int nvectors = 1) const {
    convolve(u, result, nvectors);
  }

  wdl_gauss2d* handle() const { return k_.get(); }

  double gamma;
  size_t W_, H_;
  size_t N;

 protected:
  std::shared_ptr<wdl_ctx> ctx_;
  std::shared_ptr<wdl_gauss2d> k_;
};

// Drop-in for LogSignArrayGaussian2DKernel (kernels.h:1034-1202).
class HipLogSignArrayGaussian2DKernel : public HipGaussian2DKernel {
 public:
  explicit HipLogSignArrayGaussian2DKernel(double regularization = 25., int W = 0, int H = 0)
      : HipGaussian2DKernel(regularization, W, H) {}

  // kernels.h:1073-1112
  void log_convolve(const double* log_u, double* result, int nvectors = 1) const {
    detail::check(wdl_gauss2d_log_convolve(k_.get(), log_u, result, nvectors),
                  "wdl_gauss2d_log_convolve");
  }
  void log_convolveAdjoint(const double* u, double* result, int nvectors = 1) const {
    log_convolve(u, result, nvectors);
  }
  // kernels.h:1117-1159: signArray holds the input signs and receives the output signs.
  void log_convolve_signArray(const double* log_u, unsigned char* signArray, double* result,
                              int nvectors = 1) const {
    detail::check(wdl_gauss2d_log_convolve_signarray(k_.get(), log_u, signArray, result, nvectors),
                  "wdl_gauss2d_log_convolve_signarray");
  }
  void log_convolve_signArrayAdjoint(const double* u, unsigned char* sign_array, double* result,
                                     int nvectors = 1) const {
    log_convolve_signArray(u, sign_array, result, nvectors);
  }
};

// Drop-in for Gaussian1DKernel (kernels.h:501-556): nvectors vectors of W
// values, FP64 in and out, FP32 on the GPU.
class HipGaussian1DKernel {
 public:
  // kernels.h:503-511 (default regularization 25).
  explicit HipGaussian1DKernel(double regularization = 25., int W = 0)
      : gamma(regularization), W_(W), N((size_t)W) {
    if (W > 0) {
      ctx_ = detail::shared_context();
      wdl_gauss1d* k = nullptr;
      detail::check(wdl_gauss1d_create(ctx_.get(), W, regularization, &k), "wdl_gauss1d_create");
      k_ = std::shared_ptr<wdl_gauss1d>(k, [](wdl_gauss1d* p) { wdl_gauss1d_destroy(p); });
    }
  }
  // kernels.h:530-532
  double operator()(int i, int j) const {
    return std::max(1e-250, std::exp(-(double)(i - j) * (i - j) / gamma));
  }
  // kernels.h:533-549
  void convolve(const double* u, double* result, int nvectors = 1) const {
    detail::check(wdl_gauss1d_convolve(k_.get(), u, result, nvectors), "wdl_gauss1d_convolve");
  }
  // kernels.h:550-552
  void convolveAdjoint(const double* u, double* result, int nvectors = 1) const {
    convolve(u, result, nvectors);
  }

  double gamma;
  size_t W_, N;

 private:
  std::shared_ptr<wdl_ctx> ctx_;
  std::shared_ptr<wdl_gauss1d> k_;
};

// Drop-in for Gaussian3DKernel (kernels.h:712-815): volumes of W x H x D,
// voxel (x, y, z) at z*W*H + y*W + x.
class HipGaussian3DKernel {
 public:
  // kernels.h:714-724 (default regularization 25).
  explicit HipGaussian3DKernel(double regularization = 25., int W = 0, int H = 0, int D = 0)
      : gamma(regularization), W_(W), H_(H), D_(D), N((size_t)W * H * D) {
    if (W > 0 && H > 0 && D > 0) {
      ctx_ = detail::shared_context();
      wdl_gauss3d* k = nullptr;
      detail::check(wdl_gauss3d_create(ctx_.get(), W, H, D, regularization, &k),
                    "wdl_gauss3d_create");
      k_ = std::shared_ptr<wdl_gauss3d>(k, [](wdl_gauss3d* p) { wdl_gauss3d_destroy(p); });
    }
  }
  // kernels.h:750-761: entry (i, j) of the 3-D kernel.
  double operator()(int i, int j) const {
    const int WH = (int)(W_ * H_);
    const int z0 = i / WH, r0 = i - z0 * WH, z1 = j / WH, r1 = j - z1 * WH;
    const int x0 = r0 % (int)W_, y0 = r0 / (int)W_, x1 = r1 % (int)W_, y1 = r1 / (int)W_;
    const double d2 = (double)(x0 - x1) * (x0 - x1) + (double)(y0 - y1) * (y0 - y1) +
                      (double)(z0 - z1) * (z0 - z1);
    return std::max(1e-250, std::exp(-d2 / gamma));
  }
  // kernels.h:762-807
  void convolve(const double* u, double* result, int nvectors = 1) const {
    detail::check(wdl_gauss3d_convolve(k_.get(), u, result, nvectors), "wdl_gauss3d_convolve");
  }
  // kernels.h:808-810
  void convolveAdjoint(const double* u, double* result, int nvectors = 1) const {
    convolve(u, result, nvectors);
  }

  double gamma;
  size_t W_, H_, D_, N;

 private:
  std::shared_ptr<wdl_ctx> ctx_;
  std::shared_ptr<wdl_gauss3d> k_;
};

}  // namespace wdl
