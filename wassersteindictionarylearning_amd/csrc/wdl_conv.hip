// Standard-domain separable Gaussian convolution: the drop-in for
// Gaussian2DKernel::convolve (kernels.h:669-699) and its FP32 device form.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>

#include "wdl_internal.h"

namespace wdl {

// One workgroup per image (grid-stride over images).  T is FP32 (device form)
// or FP64 (host drop-in: values are rounded to FP32 for the MFMA passes and
// widened back on store).
template <class T, bool BF3>
__global__ __launch_bounds__(kThreads, 4) void k_conv(const T* __restrict__ in,
                                                      T* __restrict__ out, int nvec, ConvGeom G,
                                                      const float* __restrict__ gtab) {
  extern __shared__ __align__(16) float smem[];
  void* tabs = smem;
  float* buf = smem + ConvPrec<BF3>::kTabBytes / 4;
  ConvPrec<BF3>::load(tabs, gtab);
  const size_t N = (size_t)G.W * G.H;
  for (int v = blockIdx.x; v < nvec; v += gridDim.x) {
    const T* src = in + (size_t)v * N;
    for_each_px(G, [&](int li, int gi, bool ok) { buf[li] = ok ? (float)src[gi] : 0.f; });
    ConvPrec<BF3>::conv(buf, tabs, G);
    T* dst = out + (size_t)v * N;
    for_each_px(G, [&](int li, int gi, bool ok) {
      if (ok) dst[gi] = (T)buf[li];
    });
    __syncthreads();
  }
}

// FP32 form with float4 loads and stores (rows of whole float4 groups).
template <bool BF3>
__global__ __launch_bounds__(kThreads, 4) void k_conv4(const float* __restrict__ in,
                                                       float* __restrict__ out, int nvec,
                                                       ConvGeom G, const float* __restrict__ gtab) {
  extern __shared__ __align__(16) float smem[];
  void* tabs = smem;
  float* buf = smem + ConvPrec<BF3>::kTabBytes / 4;
  ConvPrec<BF3>::load(tabs, gtab);
  const size_t N = (size_t)G.W * G.H;
  for (int v = blockIdx.x; v < nvec; v += gridDim.x) {
    const float* src = in + (size_t)v * N;
    px_pass<4>(
        G, [&](int gi) { return vld<4>(src + gi); },
        [&](int li, int gi, bool ok, floatx4 x) { vst<4>(buf + li, ok ? x : vsplat<4>(0.f)); });
    ConvPrec<BF3>::conv(buf, tabs, G);
    float* dst = out + (size_t)v * N;
    px_pass<4>(
        G, [&](int gi) { return 0; },
        [&](int li, int gi, bool ok, int) {
          if (ok) vst<4>(dst + gi, vld<4>(buf + li));
        });
    __syncthreads();
  }
}

// Images with a side above kMaxSide: one workgroup per 64 x 64 output tile
// (grid-stride over images x tiles), the haloed region staged in LDS.
template <class T, int VW>
__global__ __launch_bounds__(kThreads, 4) void k_conv_tile(const T* __restrict__ in,
                                                           T* __restrict__ out, int nvec,
                                                           TileGeom TG,
                                                           const float* __restrict__ gtab) {
  typedef typename Vec<VW>::T V;
  extern __shared__ __align__(16) float smem[];
  float* tabs = smem;
  float* buf = smem + kTabFloats;
  load_tables(tabs, gtab);
  const size_t N = (size_t)TG.W * TG.H;
  const int nt = TG.nx * TG.ny;
  for (long long blk = blockIdx.x; blk < (long long)nvec * nt; blk += gridDim.x) {
    const int v = (int)(blk / nt), tile = (int)(blk - (long long)v * nt);
    const int tx = tile % TG.nx, ty = tile / TG.nx;
    const T* src = in + (size_t)v * N;
    tile_region_pass<VW>(
        TG, tx, ty,
        [&](int gi) {
          if constexpr (VW == 4) return vld<4>((const float*)src + gi);
          else return (float)src[gi];
        },
        [&](int li, int, bool ok, V x) { vst<VW>(buf + li, ok ? x : vsplat<VW>(0.f)); });
    tile_conv(buf, tabs, TG);
    T* dst = out + (size_t)v * N;
    tile_inner_pass<VW>(
        TG, tx, ty, [&](int) { return 0; },
        [&](int li, int gi, bool ok, int) {
          if (!ok) return;
          if constexpr (VW == 4) vst<4>((float*)dst + gi, vld<4>(buf + li));
          else dst[gi] = (T)buf[li];
        });
    __syncthreads();
  }
}

template <class T>
static int launch_conv_tiled(const wdl_gauss2d* k, const T* in, T* out, int nvec, hipStream_t s) {
  static std::atomic<bool> attr_set{false};  // kernel objects are used from several threads
  const int lds = tile_lds_bytes(k->tg);
  if (!attr_set.load(std::memory_order_relaxed)) {
    TileGeom big = k->tg;
    big.R = kTile + 2 * kHaloMax;
    big.S = kMaxSide + 4;
    const int bytes = tile_lds_bytes(big);
    WDL_HIP(hipFuncSetAttribute((const void*)k_conv_tile<T, 1>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
    WDL_HIP(hipFuncSetAttribute((const void*)k_conv_tile<float, 4>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
    attr_set.store(true, std::memory_order_relaxed);
  }
  const long long work = (long long)nvec * k->tg.nx * k->tg.ny;
  const int grid = (int)std::min<long long>(work, 1 << 20);
  if (sizeof(T) == sizeof(float) && k->W % 4 == 0)
    hipLaunchKernelGGL((k_conv_tile<float, 4>), dim3(grid), dim3(kThreads), lds, s,
                       (const float*)in, (float*)out, nvec, k->tg, k->tab.as<float>());
  else
    hipLaunchKernelGGL((k_conv_tile<T, 1>), dim3(grid), dim3(kThreads), lds, s, in, out, nvec,
                       k->tg, k->tab.as<float>());
  WDL_HIP(hipGetLastError());
  return WDL_OK;
}

template <class T>
static int launch_conv(const wdl_gauss2d* k, const T* in, T* out, int nvec, hipStream_t s) {
  if (nvec <= 0) return WDL_OK;
  if (k->tiled) return launch_conv_tiled<T>(k, in, out, nvec, s);
  const bool bf3 = k->flags & WDL_FLAG_BF16X3;
  const ConvGeom G = k->geom(!(k->flags & WDL_FLAG_DENSE));
  static std::atomic<bool> attr_set{false};  // kernel objects are used from several threads
  if (!attr_set.load(std::memory_order_relaxed)) {
    const int img = conv_lds_bytes(make_geom(kMaxSide, kMaxSide, 0, 32)) - kTabFloats * 4;
    const void* f32[] = {(const void*)k_conv<T, false>, (const void*)k_conv4<false>};
    const void* fbf[] = {(const void*)k_conv<T, true>, (const void*)k_conv4<true>};
    for (const void* f : f32)
      WDL_HIP(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  img + ConvPrec<false>::kTabBytes));
    for (const void* f : fbf)
      WDL_HIP(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  img + ConvPrec<true>::kTabBytes));
    attr_set.store(true, std::memory_order_relaxed);
  }
  const int lds = conv_lds_bytes(G) - kTabFloats * 4 +
                  (bf3 ? ConvPrec<true>::kTabBytes : ConvPrec<false>::kTabBytes);
  const int grid = nvec;
  const bool v4 = sizeof(T) == sizeof(float) && G.W % 4 == 0;
  if (v4 && bf3)
    hipLaunchKernelGGL(k_conv4<true>, dim3(grid), dim3(kThreads), lds, s, (const float*)in,
                       (float*)out, nvec, G, k->tab.as<float>());
  else if (v4)
    hipLaunchKernelGGL(k_conv4<false>, dim3(grid), dim3(kThreads), lds, s, (const float*)in,
                       (float*)out, nvec, G, k->tab.as<float>());
  else if (bf3)
    hipLaunchKernelGGL((k_conv<T, true>), dim3(grid), dim3(kThreads), lds, s, in, out, nvec, G,
                       k->tab.as<float>());
  else
    hipLaunchKernelGGL((k_conv<T, false>), dim3(grid), dim3(kThreads), lds, s, in, out, nvec, G,
                       k->tab.as<float>());
  WDL_HIP(hipGetLastError());
  return WDL_OK;
}

int launch_conv_f32(const wdl_gauss2d* k, const float* in, float* out, int nvec, hipStream_t s) {
  return launch_conv<float>(k, in, out, nvec, s);
}

int launch_conv_f64io(const wdl_gauss2d* k, const double* in, double* out, int nvec,
                      hipStream_t s) {
  return launch_conv<double>(k, in, out, nvec, s);
}

}  // namespace wdl

wdl::ConvGeom wdl_gauss2d::geom(bool band) const {
  return wdl::make_geom(W, H, band ? D : (1 << 20), (flags & WDL_FLAG_BF16X3) ? 32 : 16);
}

extern "C" {

int wdl_gauss2d_create(wdl_ctx* ctx, int W, int H, double gamma, int flags, wdl_gauss2d** out) {
  WDL_CHECK(ctx && out, WDL_ERR_ARG, "wdl_gauss2d_create: NULL argument");
  WDL_CHECK(W > 0 && H > 0, WDL_ERR_ARG, "wdl_gauss2d_create: W and H must be positive");
  WDL_CHECK(gamma > 0 && std::isfinite(gamma), WDL_ERR_ARG,
            "wdl_gauss2d_create: gamma must be positive and finite");
  const bool tiled = wdl::needs_tiles(W, H);
  WDL_CHECK(!(tiled && (flags & (WDL_FLAG_DENSE | WDL_FLAG_BF16X3))), WDL_ERR_UNSUPPORTED,
            "wdl_gauss2d_create: WDL_FLAG_DENSE and WDL_FLAG_BF16X3 need sides of at most " +
                std::to_string(wdl::kMaxSide));
  std::vector<float> tab;
  const int D = wdl::build_table(W, H, gamma, tab);
  wdl::TileGeom tg{};
  if (tiled) {
    const int st = wdl::make_tile_geom(W, H, D, gamma, false, tg);
    if (st) return st;
  }
  WDL_HIP(hipSetDevice(ctx->device));
  wdl_gauss2d* k = new wdl_gauss2d();
  k->ctx = ctx;
  k->W = W;
  k->H = H;
  k->gamma = gamma;
  k->flags = flags;
  k->D = D;
  k->tiled = tiled;
  k->tg = tg;
  int st = k->tab.ensure(tab.size() * sizeof(float));
  if (st != WDL_OK) {
    delete k;
    return st;
  }
  hipError_t e = hipMemcpy(k->tab.p, tab.data(), tab.size() * sizeof(float), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    k->tab.release();
    delete k;
    wdl::set_error(std::string("hipMemcpy table: ") + hipGetErrorString(e));
    return WDL_ERR_HIP;
  }
  *out = k;
  return WDL_OK;
}

int wdl_gauss2d_destroy(wdl_gauss2d* k) {
  if (!k) return WDL_OK;
  (void)hipSetDevice(k->ctx->device);
  (void)hipStreamSynchronize(k->ctx->stream);
  k->tab.release();
  k->io_in.release();
  k->io_out.release();
  k->io_sign.release();
  k->io_tmp.release();
  k->io_bits.release();
  delete k;
  return WDL_OK;
}

int wdl_gauss2d_table(const wdl_gauss2d* k, double* out) {
  WDL_CHECK(k && out, WDL_ERR_ARG, "wdl_gauss2d_table: NULL argument");
  const int n = k->W > k->H ? k->W : k->H;
  for (int d = 0; d < n; ++d) out[d] = std::max(1e-250, std::exp(-(double)d * d / k->gamma));
  return WDL_OK;
}

int wdl_gauss2d_convolve(wdl_gauss2d* k, const double* u, double* out, int nvectors) {
  WDL_CHECK(k && u && out, WDL_ERR_ARG, "wdl_gauss2d_convolve: NULL argument");
  WDL_CHECK(nvectors >= 0, WDL_ERR_ARG, "wdl_gauss2d_convolve: nvectors < 0");
  if (nvectors == 0) return WDL_OK;
  WDL_HIP(hipSetDevice(k->ctx->device));
  const size_t bytes = (size_t)k->W * k->H * nvectors * sizeof(double);
  int st = k->io_in.ensure(bytes);
  if (st) return st;
  if ((st = k->io_out.ensure(bytes))) return st;
  hipStream_t s = k->ctx->stream;
  WDL_HIP(hipMemcpyAsync(k->io_in.p, u, bytes, hipMemcpyHostToDevice, s));
  if ((st = wdl::launch_conv_f64io(k, k->io_in.as<double>(), k->io_out.as<double>(), nvectors, s)))
    return st;
  WDL_HIP(hipMemcpyAsync(out, k->io_out.p, bytes, hipMemcpyDeviceToHost, s));
  WDL_HIP(hipStreamSynchronize(s));
  return WDL_OK;
}

int wdl_gauss2d_convolve_device(wdl_gauss2d* k, const float* d_u, float* d_out, int nvectors) {
  WDL_CHECK(k && d_u && d_out, WDL_ERR_ARG, "wdl_gauss2d_convolve_device: NULL argument");
  WDL_CHECK(nvectors >= 0, WDL_ERR_ARG, "wdl_gauss2d_convolve_device: nvectors < 0");
  WDL_CHECK(d_u != d_out, WDL_ERR_ARG, "wdl_gauss2d_convolve_device: in and out alias");
  WDL_HIP(hipSetDevice(k->ctx->device));
  return wdl::launch_conv_f32(k, d_u, d_out, nvectors, k->ctx->stream);
}

}  // extern "C"
