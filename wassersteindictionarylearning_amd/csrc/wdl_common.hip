// Context, errors, device buffers and the Gaussian factor table.
#include <cmath>
#include <cstring>

#include "wdl_internal.h"

namespace wdl {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

int DevBuf::ensure(size_t nbytes) {
  if (nbytes <= bytes && p) return WDL_OK;
  release();
  if (nbytes == 0) return WDL_OK;
  hipError_t e = hipMalloc(&p, nbytes);
  if (e != hipSuccess) {
    p = nullptr;
    set_error(std::string("hipMalloc(") + std::to_string(nbytes) + "): " + hipGetErrorString(e));
    return WDL_ERR_NOMEM;
  }
  bytes = nbytes;
  return WDL_OK;
}

void DevBuf::release() {
  if (p) (void)hipFree(p);
  p = nullptr;
  bytes = 0;
}

int build_table(int W, int H, double gamma, std::vector<float>& tab) {
  // kernel1d[d] = max(EPSILON, exp(-d*d/gamma)), EPSILON = 1e-250 (kernels.h:30, :636-640),
  // rounded to FP32 (1e-250 rounds to 0).
  tab.assign(kTabLen, 0.f);
  const int n = W > H ? W : H;
  int D = 0;
  for (int d = 0; d < n; ++d) {
    const double v = std::max(1e-250, std::exp(-(double)d * d / gamma));
    const float f = (float)v;
    if (kTabOff + d < kTabLen) tab[kTabOff + d] = f;
    if (kTabOff - d >= 0) tab[kTabOff - d] = f;
    if (f != 0.f) D = d;
  }
  return D;
}

static int round_up(int x, int m) { return (x + m - 1) / m * m; }

ConvGeom make_geom(int W, int H, int D, int pad) {
  ConvGeom g;
  g.W = W;
  g.H = H;
  g.Wp = round_up(W, pad);
  g.Hp = round_up(H, pad);
  // Row strides congruent to 4 (mod 64 floats): the 16 lanes of a quarter-wave
  // reading 16 consecutive rows at one column then hit distinct LDS banks.
  g.S1 = round_up(g.Wp, 64) + 4;
  g.S2 = round_up(g.Hp, 64) + 4;
  g.D = D;
  return g;
}

int make_tile_geom(int W, int H, int D, double gamma, bool logdom, TileGeom& t) {
  // Halo HB >= D: every output of the 64 x 64 tile sees all its non-zero
  // factors, and a multiple of 8, so R = 64 + 2 HB is a multiple of 16 (the
  // MFMA row blocks, the region rows of the 16 half-waves) and S / 4 is odd.
  // Positions further out would meet only factors that are 0 in FP32 --
  // also in the log domain, whose products are these factors times shifted
  // values <= 1; there a wider region would only widen the rows' shifts.
  (void)gamma;
  (void)logdom;
  const int hb = round_up(D > 0 ? D : 1, 8);
  if (hb > kHaloMax) {
    set_error("tiled convolution: gamma " + std::to_string(gamma) + " needs a halo of " +
              std::to_string(hb) + " pixels, above " + std::to_string(kHaloMax));
    return WDL_ERR_UNSUPPORTED;
  }
  t.W = W;
  t.H = H;
  t.HB = hb;
  t.R = kTile + 2 * hb;
  // R is a multiple of 16 with R / 4 even, so S / 4 = R / 4 + 1 is odd: the 16
  // rows of a quarter-wave's block start on 16 distinct groups of 4 banks (as
  // the resident strides, which are congruent to 4 mod 64), with the least LDS
  t.S = t.R + 4;
  t.D = D;
  t.nx = (W + kTile - 1) / kTile;
  t.ny = (H + kTile - 1) / kTile;
  return WDL_OK;
}

}  // namespace wdl

extern "C" {

const char* wdl_last_error(void) { return wdl::g_last_error.c_str(); }

const char* wdl_version(void) { return "wdl-gfx950 0.1"; }

int wdl_ctx_create(int device, wdl_ctx** out) {
  WDL_CHECK(out != nullptr, WDL_ERR_ARG, "wdl_ctx_create: out is NULL");
  int n = 0;
  WDL_HIP(hipGetDeviceCount(&n));
  WDL_CHECK(device >= 0 && device < n, WDL_ERR_ARG,
            "wdl_ctx_create: device " + std::to_string(device) + " of " + std::to_string(n));
  WDL_HIP(hipSetDevice(device));
  wdl_ctx* c = new wdl_ctx();
  c->device = device;
  hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete c;
    wdl::set_error(std::string("hipStreamCreate: ") + hipGetErrorString(e));
    return WDL_ERR_HIP;
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->num_cus = prop.multiProcessorCount;
  int st = wdl::noalloc_scratch_init(device);
  if (st) {
    (void)hipStreamDestroy(c->stream);
    delete c;
    return st;
  }
  *out = c;
  return WDL_OK;
}

int wdl_ctx_destroy(wdl_ctx* ctx) {
  if (!ctx) return WDL_OK;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) {
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipStreamDestroy(ctx->stream);
  }
  delete ctx;
  return WDL_OK;
}

void* wdl_ctx_stream(wdl_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

int wdl_ctx_query(wdl_ctx* ctx) {
  WDL_CHECK(ctx != nullptr, WDL_ERR_ARG, "wdl_ctx_query: ctx is NULL");
  const hipError_t e = hipStreamQuery(ctx->stream);
  if (e == hipSuccess) return 0;
  if (e == hipErrorNotReady) return 1;
  wdl::set_error(std::string("hipStreamQuery: ") + hipGetErrorString(e));
  return WDL_ERR_HIP;
}

int wdl_ctx_synchronize(wdl_ctx* ctx) {
  WDL_CHECK(ctx != nullptr, WDL_ERR_ARG, "wdl_ctx_synchronize: ctx is NULL");
  WDL_HIP(hipStreamSynchronize(ctx->stream));
  return WDL_OK;
}

}  // extern "C"
