// 1-D and 3-D separable Gaussian convolutions: the drop-ins for
// Gaussian1DKernel::convolve (kernels.h:533-552) and
// Gaussian3DKernel::convolve (kernels.h:762-810), and their FP32 device forms.
//
// Both rest on one line pass: for a tensor viewed as [outer][len][inner], every
// line (o, i) of `len` values at stride `inner` is convolved with the 1-D
// Gaussian, OUT[o][j][i] = sum_k tab[|j - k|] IN[o][k][i].  A workgroup stages
// 64 lines (16 when a line is longer than kMaxSide) in LDS and runs the same
// MFMA pass as the 2-D kernels (rows_gemm, band skipping included) in blocks of
// up to 256 outputs.  1-D: outer = nvectors, inner = 1.  3-D: the 2-D
// convolution of every z-slice (x then y, :768-789), then the line pass along z
// (outer = nvectors, len = D, inner = W*H, :791-803).
#include <atomic>
#include <algorithm>
#include <cmath>
#include <vector>

#include "wdl_internal.h"

namespace wdl {

constexpr int kLineMaxLen = 2048;
constexpr int kLineBlock = kMaxUnits * kWaves * 16;  // outputs per rows_gemm call

struct LineGeom {
  int len;    // line length
  int Lp;     // padded to 16
  int lpw;    // lines per workgroup (a multiple of 16)
  int S;      // LDS stride of a staged line (floats), congruent to 4 mod 64
  int D;      // FP32 band half-width
};

static LineGeom make_line_geom(int len, int D) {
  LineGeom g;
  g.len = len;
  g.Lp = (len + 15) / 16 * 16;
  g.lpw = g.Lp <= kMaxSide ? 64 : 16;
  g.S = (g.Lp + 63) / 64 * 64 + 4;
  g.D = D;
  return g;
}

static int line_lds_bytes(const LineGeom& g) { return (kTabFloats + g.lpw * g.S) * 4; }

template <class T>
__global__ __launch_bounds__(kThreads) void k_line(const T* __restrict__ in, T* __restrict__ out,
                                                   long long outer, long long inner, LineGeom g,
                                                   const float* __restrict__ gtab) {
  extern __shared__ __align__(16) float smem[];
  float* tabs = smem;
  float* buf = smem + kTabFloats;
  load_tables(tabs, gtab);
  const long long nlines = outer * inner;
  const long long ntiles = (nlines + g.lpw - 1) / g.lpw;
  const int lane = threadIdx.x & 63, wave = wave_id();
  const int c = lane & 15, q = lane >> 4;
  for (long long tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const long long l0 = tile * g.lpw;
    // Stage line r (= l0 + r), position k at buf[r * S + k]; consecutive threads
    // take consecutive addresses: along k when inner == 1, across lines else.
    const int tot = g.lpw * g.Lp;
    for (int e = threadIdx.x; e < tot; e += kThreads) {
      int r, k;
      if (inner == 1) {
        r = e / g.Lp;
        k = e - r * g.Lp;
      } else {
        k = e / g.lpw;
        r = e - k * g.lpw;
      }
      const long long l = l0 + r;
      float v = 0.f;
      if (l < nlines && k < g.len) {
        const long long o = l / inner, i = l - o * inner;
        v = (float)in[(o * g.len + k) * inner + i];
      }
      buf[r * g.S + k] = v;
    }
    __syncthreads();
    for (int ob = 0; ob < g.Lp; ob += kLineBlock) {
      floatx4 acc[kMaxUnits][4];
      const int nout = min(kLineBlock, g.Lp - ob);
      rows_gemm(buf, g.S, g.lpw, g.Lp, nout, ob, tabs, g.D, acc);
      // acc[uu][rb][j]: line 64 rg + 16 rb + 4 q + j, output ob + 16 cb + c
      const int ncb = nout >> 4;
      const int nunits = ncb * ((g.lpw + 63) >> 6);
#pragma unroll
      for (int uu = 0; uu < kMaxUnits; ++uu) {
        const int u = wave + uu * kWaves;
        if (u >= nunits) continue;
        const int cb = u % ncb, rg = u / ncb;
        const int col = ob + 16 * cb + c;
        const int nrb = min(4, (g.lpw - 64 * rg) >> 4);
#pragma unroll
        for (int rb = 0; rb < 4; ++rb) {
          if (rb >= nrb) continue;
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            const long long l = l0 + 64 * rg + 16 * rb + 4 * q + j;
            if (l < nlines && col < g.len) {
              const long long o = l / inner, i = l - o * inner;
              out[(o * g.len + col) * inner + i] = (T)acc[uu][rb][j];
            }
          }
        }
      }
    }
    __syncthreads();
  }
}

static int line_set_lds_attr() {
  static std::atomic<bool> done{false};  // kernel objects are created from several threads
  if (done.load(std::memory_order_relaxed)) return WDL_OK;
  const int b = line_lds_bytes(make_line_geom(kLineMaxLen, 0));
  WDL_HIP(hipFuncSetAttribute((const void*)k_line<float>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, b));
  WDL_HIP(hipFuncSetAttribute((const void*)k_line<double>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, b));
  done.store(true, std::memory_order_relaxed);
  return WDL_OK;
}

template <class T>
static int launch_line(const T* in, T* out, long long outer, long long inner, const LineGeom& g,
                       const float* tab, hipStream_t s) {
  const long long nlines = outer * inner;
  if (nlines == 0) return WDL_OK;
  const long long ntiles = (nlines + g.lpw - 1) / g.lpw;
  const int grid = (int)std::min<long long>(ntiles, 1 << 16);
  hipLaunchKernelGGL(k_line<T>, dim3(grid), dim3(kThreads), line_lds_bytes(g), s, in, out, outer,
                     inner, g, tab);
  WDL_HIP(hipGetLastError());
  return WDL_OK;
}

// The FP32 factor table of a line of `len` values and its band half-width;
// checks that the band fits the table (every line up to kMaxSide does; longer
// lines need D <= 114, i.e. gamma up to ~125).
static int line_table(int len, double gamma, std::vector<float>& tab, int& D, const char* who) {
  D = build_table(len, 1, gamma, tab);
  WDL_CHECK(len <= kMaxSide || D <= 114, WDL_ERR_UNSUPPORTED,
            std::string(who) + ": lines above " + std::to_string(kMaxSide) +
                " values need gamma small enough for a band of at most 114 (here " +
                std::to_string(D) + ")");
  return WDL_OK;
}

}  // namespace wdl

struct wdl_gauss1d {
  wdl_ctx* ctx = nullptr;
  int W = 0;
  double gamma = 0;
  wdl::LineGeom g{};
  wdl::DevBuf tab, io_in, io_out;
};

struct wdl_gauss3d {
  wdl_ctx* ctx = nullptr;
  int W = 0, H = 0, D = 0;
  double gamma = 0;
  wdl_gauss2d* slice = nullptr;  // x and y passes of every z-slice
  wdl::LineGeom gz{};
  wdl::DevBuf tab, io_in, io_out, tmp;
};

extern "C" {

int wdl_gauss1d_create(wdl_ctx* ctx, int W, double gamma, wdl_gauss1d** out) {
  WDL_CHECK(ctx && out, WDL_ERR_ARG, "wdl_gauss1d_create: NULL argument");
  WDL_CHECK(W > 0 && W <= wdl::kLineMaxLen, WDL_ERR_ARG,
            "wdl_gauss1d_create: W must be in 1.." + std::to_string(wdl::kLineMaxLen));
  WDL_CHECK(gamma > 0 && std::isfinite(gamma), WDL_ERR_ARG,
            "wdl_gauss1d_create: gamma must be positive and finite");
  std::vector<float> tab;
  int D = 0, st;
  if ((st = wdl::line_table(W, gamma, tab, D, "wdl_gauss1d_create"))) return st;
  WDL_HIP(hipSetDevice(ctx->device));
  if ((st = wdl::line_set_lds_attr())) return st;
  wdl_gauss1d* k = new wdl_gauss1d();
  k->ctx = ctx;
  k->W = W;
  k->gamma = gamma;
  k->g = wdl::make_line_geom(W, D);
  if ((st = k->tab.ensure(tab.size() * sizeof(float)))) {
    delete k;
    return st;
  }
  if (hipMemcpy(k->tab.p, tab.data(), tab.size() * sizeof(float), hipMemcpyHostToDevice) !=
      hipSuccess) {
    k->tab.release();
    delete k;
    wdl::set_error("wdl_gauss1d_create: copy of the factor table failed");
    return WDL_ERR_HIP;
  }
  *out = k;
  return WDL_OK;
}

int wdl_gauss1d_destroy(wdl_gauss1d* k) {
  if (!k) return WDL_OK;
  (void)hipSetDevice(k->ctx->device);
  (void)hipStreamSynchronize(k->ctx->stream);
  k->tab.release();
  k->io_in.release();
  k->io_out.release();
  delete k;
  return WDL_OK;
}

int wdl_gauss1d_convolve(wdl_gauss1d* k, const double* u, double* out, int nvectors) {
  WDL_CHECK(k && u && out, WDL_ERR_ARG, "wdl_gauss1d_convolve: NULL argument");
  WDL_CHECK(nvectors >= 0, WDL_ERR_ARG, "wdl_gauss1d_convolve: nvectors < 0");
  if (nvectors == 0) return WDL_OK;
  WDL_HIP(hipSetDevice(k->ctx->device));
  const size_t bytes = (size_t)k->W * nvectors * sizeof(double);
  int st;
  if ((st = k->io_in.ensure(bytes)) || (st = k->io_out.ensure(bytes))) return st;
  hipStream_t s = k->ctx->stream;
  WDL_HIP(hipMemcpyAsync(k->io_in.p, u, bytes, hipMemcpyHostToDevice, s));
  if ((st = wdl::launch_line(k->io_in.as<double>(), k->io_out.as<double>(), nvectors, 1, k->g,
                             k->tab.as<float>(), s)))
    return st;
  WDL_HIP(hipMemcpyAsync(out, k->io_out.p, bytes, hipMemcpyDeviceToHost, s));
  WDL_HIP(hipStreamSynchronize(s));
  return WDL_OK;
}

int wdl_gauss1d_convolve_device(wdl_gauss1d* k, const float* d_u, float* d_out, int nvectors) {
  WDL_CHECK(k && d_u && d_out, WDL_ERR_ARG, "wdl_gauss1d_convolve_device: NULL argument");
  WDL_CHECK(nvectors >= 0, WDL_ERR_ARG, "wdl_gauss1d_convolve_device: nvectors < 0");
  WDL_CHECK(d_u != d_out, WDL_ERR_ARG, "wdl_gauss1d_convolve_device: in and out alias");
  WDL_HIP(hipSetDevice(k->ctx->device));
  return wdl::launch_line(d_u, d_out, nvectors, 1, k->g, k->tab.as<float>(), k->ctx->stream);
}

int wdl_gauss3d_create(wdl_ctx* ctx, int W, int H, int D, double gamma, wdl_gauss3d** out) {
  WDL_CHECK(ctx && out, WDL_ERR_ARG, "wdl_gauss3d_create: NULL argument");
  WDL_CHECK(W > 0 && H > 0 && D > 0, WDL_ERR_ARG, "wdl_gauss3d_create: W, H, D must be positive");
  WDL_CHECK(D <= wdl::kLineMaxLen, WDL_ERR_UNSUPPORTED,
            "wdl_gauss3d_create: D must be at most " + std::to_string(wdl::kLineMaxLen));
  WDL_CHECK(gamma > 0 && std::isfinite(gamma), WDL_ERR_ARG,
            "wdl_gauss3d_create: gamma must be positive and finite");
  std::vector<float> tab;
  int Dz = 0, st;
  if ((st = wdl::line_table(D, gamma, tab, Dz, "wdl_gauss3d_create"))) return st;
  wdl_gauss2d* slice = nullptr;
  if ((st = wdl_gauss2d_create(ctx, W, H, gamma, 0, &slice))) return st;
  if ((st = wdl::line_set_lds_attr())) {
    wdl_gauss2d_destroy(slice);
    return st;
  }
  wdl_gauss3d* k = new wdl_gauss3d();
  k->ctx = ctx;
  k->W = W;
  k->H = H;
  k->D = D;
  k->gamma = gamma;
  k->slice = slice;
  k->gz = wdl::make_line_geom(D, Dz);
  if ((st = k->tab.ensure(tab.size() * sizeof(float))) ||
      hipMemcpy(k->tab.p, tab.data(), tab.size() * sizeof(float), hipMemcpyHostToDevice) !=
          hipSuccess) {
    wdl_gauss3d_destroy(k);
    if (st) return st;
    wdl::set_error("wdl_gauss3d_create: copy of the factor table failed");
    return WDL_ERR_HIP;
  }
  *out = k;
  return WDL_OK;
}

int wdl_gauss3d_destroy(wdl_gauss3d* k) {
  if (!k) return WDL_OK;
  (void)hipSetDevice(k->ctx->device);
  (void)hipStreamSynchronize(k->ctx->stream);
  wdl_gauss2d_destroy(k->slice);
  k->tab.release();
  k->io_in.release();
  k->io_out.release();
  k->tmp.release();
  delete k;
  return WDL_OK;
}

int wdl_gauss3d_convolve(wdl_gauss3d* k, const double* u, double* out, int nvectors) {
  WDL_CHECK(k && u && out, WDL_ERR_ARG, "wdl_gauss3d_convolve: NULL argument");
  WDL_CHECK(nvectors >= 0, WDL_ERR_ARG, "wdl_gauss3d_convolve: nvectors < 0");
  if (nvectors == 0) return WDL_OK;
  WDL_HIP(hipSetDevice(k->ctx->device));
  const long long WH = (long long)k->W * k->H;
  const size_t bytes = (size_t)WH * k->D * nvectors * sizeof(double);
  int st;
  if ((st = k->io_in.ensure(bytes)) || (st = k->io_out.ensure(bytes)) || (st = k->tmp.ensure(bytes)))
    return st;
  hipStream_t s = k->ctx->stream;
  WDL_HIP(hipMemcpyAsync(k->io_in.p, u, bytes, hipMemcpyHostToDevice, s));
  if ((st = wdl::launch_conv_f64io(k->slice, k->io_in.as<double>(), k->tmp.as<double>(),
                                   k->D * nvectors, s)))
    return st;
  if ((st = wdl::launch_line(k->tmp.as<double>(), k->io_out.as<double>(), nvectors, WH, k->gz,
                             k->tab.as<float>(), s)))
    return st;
  WDL_HIP(hipMemcpyAsync(out, k->io_out.p, bytes, hipMemcpyDeviceToHost, s));
  WDL_HIP(hipStreamSynchronize(s));
  return WDL_OK;
}

int wdl_gauss3d_convolve_device(wdl_gauss3d* k, const float* d_u, float* d_out, int nvectors) {
  WDL_CHECK(k && d_u && d_out, WDL_ERR_ARG, "wdl_gauss3d_convolve_device: NULL argument");
  WDL_CHECK(nvectors >= 0, WDL_ERR_ARG, "wdl_gauss3d_convolve_device: nvectors < 0");
  WDL_CHECK(d_u != d_out, WDL_ERR_ARG, "wdl_gauss3d_convolve_device: in and out alias");
  if (nvectors == 0) return WDL_OK;
  WDL_HIP(hipSetDevice(k->ctx->device));
  const long long WH = (long long)k->W * k->H;
  int st;
  if ((st = k->tmp.ensure((size_t)WH * k->D * nvectors * sizeof(float)))) return st;
  hipStream_t s = k->ctx->stream;
  if ((st = wdl::launch_conv_f32(k->slice, d_u, k->tmp.as<float>(), k->D * nvectors, s)))
    return st;
  return wdl::launch_line(k->tmp.as<float>(), d_out, nvectors, WH, k->gz, k->tab.as<float>(), s);
}

}  // extern "C"
