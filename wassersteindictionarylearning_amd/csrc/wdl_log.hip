// Log-domain separable Gaussian convolution: the drop-in for
// LogSignArrayGaussian2DKernel::log_convolve / log_convolve_signArray
// (kernels.h:1073-1162), in FP64 on the GPU.
//
// Each 1-D pass computes, for every output j of a line x,
//     m_j = max(-1e99, max_k(-(j-k)^2/gamma + x_k))
//     s_j = sum_k sign_k * exp(-(j-k)^2/gamma + x_k - m_j)
//     out_j = log|s_j| + m_j,   sign_out_j = (s_j >= 0)
// with k ascending, as the reference does; one workgroup per line, the line
// staged in LDS, one thread per output.  Pass 1 runs along x and writes its
// result transposed (as the reference's tmp[i + j*H]); pass 2 runs along y on
// that and writes back in image order.  Output signs go to a byte per value
// and are packed LSB-first into the caller's sign array (signArray.h:8-17)
// afterwards, so no two threads write one byte.
#include <atomic>
#include <cmath>
#include <mutex>
#include <string>

#include "wdl_internal.h"

namespace wdl {

constexpr int kLineThreads = 256;

// in: nvec images of `rows` lines of length `len` (line-major).  out: the
// transposed layout (len x rows per image).  Input signs either from packed
// bits (bit index = v*rows*len + r*len + k) or from bytes; none means positive.
template <bool SIGNED, bool BITS_IN>
__global__ __launch_bounds__(kLineThreads) void k_lse_lines(
    const double* __restrict__ in, const unsigned char* __restrict__ sgn_in, double* __restrict__ out,
    unsigned char* __restrict__ sgn_out, int rows, int len, int nvec, double gamma) {
  extern __shared__ __align__(16) double line[];
  double* x = line;
  float* sg = reinterpret_cast<float*>(line + len);
  const long long L = (long long)rows * len;
  for (long long blk = blockIdx.x; blk < (long long)nvec * rows; blk += gridDim.x) {
    const int v = (int)(blk / rows), r = (int)(blk - (long long)v * rows);
    const long long base = (long long)v * L + (long long)r * len;
    for (int k = threadIdx.x; k < len; k += kLineThreads) {
      x[k] = in[base + k];
      if (SIGNED) {
        bool pos;
        if (BITS_IN) {
          const long long bit = base + k;
          pos = (sgn_in[bit >> 3] >> (bit & 7)) & 1;
        } else {
          pos = sgn_in[base + k] != 0;
        }
        sg[k] = pos ? 1.f : -1.f;  // getSign (signArray.h:19-22)
      }
    }
    __syncthreads();
    for (int j = threadIdx.x; j < len; j += kLineThreads) {
      double m = -1e99;
      for (int k = 0; k < len; ++k) {
        const double t = -(double)((j - k) * (j - k)) / gamma + x[k];
        if (m < t) m = t;  // std::max(m, t)
      }
      double s = 0;
      for (int k = 0; k < len; ++k) {
        const double e = exp(-(double)((j - k) * (j - k)) / gamma + x[k] - m);
        s += SIGNED ? (double)sg[k] * e : e;
      }
      const long long o = (long long)v * L + (long long)j * rows + r;  // transposed
      out[o] = log(SIGNED ? fabs(s) : s) + m;
      if (SIGNED) sgn_out[o] = s >= 0 ? 1 : 0;
    }
    __syncthreads();
  }
}

// Packs one byte per value into bits (LSB first); bits past n keep their value.
__global__ void k_pack_signs(const unsigned char* __restrict__ bytes, unsigned char* __restrict__ bits,
                             long long n) {
  const long long nb = (n + 7) >> 3;
  for (long long b = (long long)blockIdx.x * blockDim.x + threadIdx.x; b < nb;
       b += (long long)gridDim.x * blockDim.x) {
    unsigned char byte = bits[b];
    for (int t = 0; t < 8; ++t) {
      const long long i = b * 8 + t;
      if (i < n) byte = bytes[i] ? (byte | (1u << t)) : (byte & ~(1u << t));
    }
    bits[b] = byte;
  }
}

// Unpacks bits (LSB first) into one byte per value.
__global__ void k_unpack_signs(const unsigned char* __restrict__ bits,
                               unsigned char* __restrict__ bytes, long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    bytes[i] = (bits[i >> 3] >> (i & 7)) & 1;
}

static int lse_launch_bytes(int len) { return len * (int)(sizeof(double) + sizeof(float)); }

// Two passes over nvec images; d_bits == nullptr for the unsigned form.
// Scratch: tmp holds the transposed FP64 intermediate (nvec*W*H values), sg
// the sign bytes (3*nvec*W*H: input, intermediate and output signs).
static int log_convolve_core(const wdl_gauss2d* k, const double* d_in, double* d_out,
                             unsigned char* d_bits, int nvec, double* tmp, unsigned char* sg,
                             hipStream_t st) {
  const long long n = (long long)k->W * k->H * nvec;
  const bool sgn = d_bits != nullptr;
  unsigned char* b1 = sg;
  unsigned char* b2 = sg ? sg + n : nullptr;
  const int W = k->W, H = k->H;
  const int g1 = (int)std::min<long long>((long long)nvec * H, 65535);
  const int g2 = (int)std::min<long long>((long long)nvec * W, 65535);
  int rc;
  if (k->flags & WDL_FLAG_EXACT_LSE) {  // exact-shift passes, FP32 sums (wdl_xlse.hip)
    WDL_CHECK(std::max(W, H) <= 1024, WDL_ERR_UNSUPPORTED,
              "log_convolve: WDL_FLAG_EXACT_LSE takes sides up to 1024");
    unsigned char* b0 = sg ? sg + 2 * n : nullptr;
    const int gp = (int)std::min<long long>((n + 255) / 256, 8192);
    if (sgn) hipLaunchKernelGGL(k_unpack_signs, dim3(gp), dim3(256), 0, st, d_bits, b0, n);
    if ((rc = xlse_convolve(d_in, sgn ? b0 : nullptr, d_out, sgn ? b2 : nullptr, tmp,
                            sgn ? b1 : nullptr, W, H, nvec, k->gamma, st)))
      return rc;
    if (sgn) {
      const int gq = (int)std::min<long long>(((n + 7) / 8 + 255) / 256, 4096);
      hipLaunchKernelGGL(k_pack_signs, dim3(gq), dim3(256), 0, st, b2, d_bits, n);
    }
    WDL_HIP(hipGetLastError());
    return WDL_OK;
  }
  if (sgn) {
    hipLaunchKernelGGL((k_lse_lines<true, true>), dim3(g1), dim3(kLineThreads), lse_launch_bytes(W),
                       st, d_in, d_bits, tmp, b1, H, W, nvec, k->gamma);
    hipLaunchKernelGGL((k_lse_lines<true, false>), dim3(g2), dim3(kLineThreads),
                       lse_launch_bytes(H), st, tmp, b1, d_out, b2, W, H, nvec, k->gamma);
    const int gp = (int)std::min<long long>(((n + 7) / 8 + 255) / 256, 4096);
    hipLaunchKernelGGL(k_pack_signs, dim3(gp), dim3(256), 0, st, b2, d_bits, n);
  } else {
    hipLaunchKernelGGL((k_lse_lines<false, false>), dim3(g1), dim3(kLineThreads),
                       lse_launch_bytes(W), st, d_in, nullptr, tmp, nullptr, H, W, nvec, k->gamma);
    hipLaunchKernelGGL((k_lse_lines<false, false>), dim3(g2), dim3(kLineThreads),
                       lse_launch_bytes(H), st, tmp, nullptr, d_out, nullptr, W, H, nvec, k->gamma);
  }
  WDL_HIP(hipGetLastError());
  return WDL_OK;
}

// The kernel object's own scratch, grown on demand (hipMalloc).
static int log_convolve_dev(wdl_gauss2d* k, const double* d_in, double* d_out, unsigned char* d_bits,
                            int nvec, hipStream_t st) {
  const long long n = (long long)k->W * k->H * nvec;
  int rc;
  if ((rc = k->io_tmp.ensure(n * sizeof(double)))) return rc;
  if (d_bits && (rc = k->io_sign.ensure(3 * n))) return rc;
  return log_convolve_core(k, d_in, d_out, d_bits, nvec, k->io_tmp.as<double>(),
                           d_bits ? k->io_sign.as<unsigned char>() : nullptr, st);
}

// ---- allocation-free form ------------------------------------------------------
// Scratch of wdl_gauss2d_log_convolve_noalloc: the FP64 intermediate of up to
// WDL_NOALLOC_MAX_VALUES values, then 3 sign bytes per value.  One copy per
// GPU, held while the library is loaded; every call on a GPU shares it.
constexpr long long kNoAllocValues = WDL_NOALLOC_MAX_VALUES;
constexpr size_t kNoAllocBytes = (size_t)kNoAllocValues * (sizeof(double) + 3);
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

constexpr int kMaxDevices = 64;
static std::atomic<unsigned char*> g_scratch[kMaxDevices];
static std::mutex g_scratch_mu;

// The device address of `scratch` on `device`, which must be current.  Looked
// up once per device (from wdl_ctx_create); host code uses only this address,
// never the array's name, which on the host is a stand-in and no device address.
int noalloc_scratch_init(int device) {
  WDL_CHECK(device >= 0 && device < kMaxDevices, WDL_ERR_UNSUPPORTED,
            "noalloc scratch: device index above " + std::to_string(kMaxDevices - 1));
  if (g_scratch[device].load(std::memory_order_acquire)) return WDL_OK;
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (g_scratch[device].load(std::memory_order_relaxed)) return WDL_OK;
  void* p = nullptr;
  WDL_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  WDL_CHECK(p != nullptr, WDL_ERR_HIP, "noalloc scratch: hipGetSymbolAddress gave NULL");
  g_scratch[device].store(static_cast<unsigned char*>(p), std::memory_order_release);
  return WDL_OK;
}

}  // namespace wdl

extern "C" {

int wdl_gauss2d_log_convolve(wdl_gauss2d* k, const double* log_u, double* out, int nvectors) {
  WDL_CHECK(k && log_u && out, WDL_ERR_ARG, "wdl_gauss2d_log_convolve: NULL argument");
  WDL_CHECK(nvectors >= 0, WDL_ERR_ARG, "wdl_gauss2d_log_convolve: nvectors < 0");
  if (nvectors == 0) return WDL_OK;
  WDL_HIP(hipSetDevice(k->ctx->device));
  const size_t bytes = (size_t)k->W * k->H * nvectors * sizeof(double);
  int st;
  if ((st = k->io_in.ensure(bytes)) || (st = k->io_out.ensure(bytes))) return st;
  hipStream_t s = k->ctx->stream;
  WDL_HIP(hipMemcpyAsync(k->io_in.p, log_u, bytes, hipMemcpyHostToDevice, s));
  if ((st = wdl::log_convolve_dev(k, k->io_in.as<double>(), k->io_out.as<double>(), nullptr,
                                  nvectors, s)))
    return st;
  WDL_HIP(hipMemcpyAsync(out, k->io_out.p, bytes, hipMemcpyDeviceToHost, s));
  WDL_HIP(hipStreamSynchronize(s));
  return WDL_OK;
}

int wdl_gauss2d_log_convolve_signarray(wdl_gauss2d* k, const double* log_u, unsigned char* signs,
                                       double* out, int nvectors) {
  WDL_CHECK(k && log_u && signs && out, WDL_ERR_ARG,
            "wdl_gauss2d_log_convolve_signarray: NULL argument");
  WDL_CHECK(nvectors >= 0, WDL_ERR_ARG, "wdl_gauss2d_log_convolve_signarray: nvectors < 0");
  if (nvectors == 0) return WDL_OK;
  WDL_HIP(hipSetDevice(k->ctx->device));
  const long long n = (long long)k->W * k->H * nvectors;
  const size_t bytes = (size_t)n * sizeof(double);
  const size_t nbits = (size_t)(n + 7) / 8;
  int st;
  if ((st = k->io_in.ensure(bytes)) || (st = k->io_out.ensure(bytes)) ||
      (st = k->io_bits.ensure(nbits)))
    return st;
  hipStream_t s = k->ctx->stream;
  WDL_HIP(hipMemcpyAsync(k->io_in.p, log_u, bytes, hipMemcpyHostToDevice, s));
  WDL_HIP(hipMemcpyAsync(k->io_bits.p, signs, nbits, hipMemcpyHostToDevice, s));
  if ((st = wdl::log_convolve_dev(k, k->io_in.as<double>(), k->io_out.as<double>(),
                                  k->io_bits.as<unsigned char>(), nvectors, s)))
    return st;
  WDL_HIP(hipMemcpyAsync(out, k->io_out.p, bytes, hipMemcpyDeviceToHost, s));
  WDL_HIP(hipMemcpyAsync(signs, k->io_bits.p, nbits, hipMemcpyDeviceToHost, s));
  WDL_HIP(hipStreamSynchronize(s));
  return WDL_OK;
}

int wdl_gauss2d_log_convolve_noalloc(wdl_gauss2d* k, const double* d_log_u, unsigned char* d_signs,
                                     double* d_out, int nvectors) {
  WDL_CHECK(k && d_log_u && d_out, WDL_ERR_ARG, "wdl_gauss2d_log_convolve_noalloc: NULL argument");
  WDL_CHECK(nvectors >= 0, WDL_ERR_ARG, "wdl_gauss2d_log_convolve_noalloc: nvectors < 0");
  const long long n = (long long)k->W * k->H * nvectors;
  WDL_CHECK(n <= wdl::kNoAllocValues, WDL_ERR_ARG,
            "wdl_gauss2d_log_convolve_noalloc: nvectors*W*H = " + std::to_string(n) +
                " exceeds WDL_NOALLOC_MAX_VALUES = " + std::to_string(wdl::kNoAllocValues));
  if (nvectors == 0) return WDL_OK;
  const int dev = k->ctx->device;
  WDL_HIP(hipSetDevice(dev));
  int st;
  if ((st = wdl::noalloc_scratch_init(dev))) return st;  // set at wdl_ctx_create already
  unsigned char* base = wdl::g_scratch[dev].load(std::memory_order_acquire);
  double* tmp = reinterpret_cast<double*>(base);
  unsigned char* sg = base + (size_t)wdl::kNoAllocValues * sizeof(double);
  return wdl::log_convolve_core(k, d_log_u, d_out, d_signs, nvectors, tmp, d_signs ? sg : nullptr,
                                k->ctx->stream);
}

}  // extern "C"
