// Log-domain separable Gaussian convolution (LogSignArrayGaussian2DKernel,
// kernels.h:1034-1202).
#include "wdl_internal.h"

extern "C" {

int wdl_gauss2d_log_convolve(wdl_gauss2d* k, const double* log_u, double* out, int nvectors) {
  (void)k; (void)log_u; (void)out; (void)nvectors;
  wdl::set_error("wdl_gauss2d_log_convolve: not available in this build");
  return WDL_ERR_UNSUPPORTED;
}

int wdl_gauss2d_log_convolve_signarray(wdl_gauss2d* k, const double* log_u, unsigned char* signs,
                                       double* out, int nvectors) {
  (void)k; (void)log_u; (void)signs; (void)out; (void)nvectors;
  wdl::set_error("wdl_gauss2d_log_convolve_signarray: not available in this build");
  return WDL_ERR_UNSUPPORTED;
}

}  // extern "C"
