// Host-side internals shared by the HIP translation units of libwdl.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/wdl_c.h"
#include "wdl_conv_core.h"
#include "wdl_tile.h"

namespace wdl {

// Thread-local message behind wdl_last_error().
void set_error(const std::string& msg);

#define WDL_HIP(expr)                                                                   \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess) {                                                             \
      ::wdl::set_error(std::string(#expr) + ": " + hipGetErrorString(e_));              \
      return WDL_ERR_HIP;                                                               \
    }                                                                                   \
  } while (0)

#define WDL_CHECK(cond, code, msg)   \
  do {                               \
    if (!(cond)) {                   \
      ::wdl::set_error(msg);         \
      return (code);                 \
    }                                \
  } while (0)

// Device buffer owned by the library (freed in the owner's destructor).
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  int ensure(size_t nbytes);   // grows (never shrinks); returns status
  void release();
  template <class T> T* as() const { return static_cast<T*>(p); }
};

// FP32 table for the LDS copies: tab[kTabOff + d] = (float)max(1e-250, exp(-d^2/gamma))
// for |d| < max(W,H), 0 elsewhere (kernels.h:636-640).  Returns D, the largest
// |d| whose FP32 factor is non-zero.
int build_table(int W, int H, double gamma, std::vector<float>& tab);

// pad: padded sides are multiples of it (16 for the FP32 passes, 32 for the
// split-BF16x3 passes, whose contraction chunks are 32 wide).
ConvGeom make_geom(int W, int H, int D, int pad = 16);

// Geometry of the tiled convolution for band half-width D; the log domain
// (logdom) widens the halo until the factors beyond it are below e^-264.
// Returns WDL_ERR_UNSUPPORTED when the halo would exceed kHaloMax.
int make_tile_geom(int W, int H, int D, double gamma, bool logdom, TileGeom& t);

// Exact-shift log-domain convolution of nimg W x H images (wdl_xlse.hip):
// signs one byte per value (nullptr: all non-negative); tmp / stmp hold the
// transposed intermediate.
int xlse_convolve(const double* in, const unsigned char* sin, double* out, unsigned char* sout,
                  double* tmp, unsigned char* stmp, int W, int H, long long nimg, double gamma,
                  hipStream_t st);

// Looks up, once per device, the device address of the allocation-free entry
// point's static scratch (wdl_log.hip); the device must be current.
int noalloc_scratch_init(int device);

// Images with a side above kMaxSide take the tiled kernels.
inline bool needs_tiles(int W, int H) { return W > kMaxSide || H > kMaxSide; }

}  // namespace wdl

struct wdl_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  int num_cus = 0;
};

struct wdl_gauss2d {
  wdl_ctx* ctx = nullptr;
  int W = 0, H = 0;
  double gamma = 0;
  int D = 0;              // FP32 band half-width
  int flags = 0;
  bool tiled = false;     // a side above kMaxSide: 64 x 64 tiles with a halo
  wdl::TileGeom tg;
  wdl::DevBuf tab;        // kTabLen floats
  wdl::DevBuf io_in, io_out;   // staging for the host-pointer entry points
  wdl::DevBuf io_tmp;          // log domain: the transposed pass-1 result (FP64)
  wdl::DevBuf io_sign;         // log domain: one byte per value, passes 1 and 2
  wdl::DevBuf io_bits;         // log domain: the caller's packed sign array
  wdl::ConvGeom geom(bool band) const;
};

namespace wdl {

// Launchers (stream-ordered, no synchronisation).
int launch_conv_f32(const wdl_gauss2d* k, const float* in, float* out, int nvec, hipStream_t s);
int launch_conv_f64io(const wdl_gauss2d* k, const double* in, double* out, int nvec,
                      hipStream_t s);

}  // namespace wdl
