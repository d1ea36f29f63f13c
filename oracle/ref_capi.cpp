// Test-only C entry points over the reference's own templates.
//
// This file is compiled together with the unmodified reference headers found
// under /root/reference/cpp (see oracle/Makefile) into oracle/_ref/*.so.  It is
// the checker's checker: tests use it to pin the numpy oracle and to write the
// golden fixtures, and bench.py times it as the CPU baseline.  Nothing in the
// product links it.
//
// Two shared objects are built from this one file:
//   libwdlref.so      standard domain   (inverseWasserstein.h:544-674, :986-1145)
//   libwdlref_log.so  -DCOMPUTE_BARYCENTER_LOG (inverseWasserstein.h:678-814 via :1067)
// because evaluate_both picks its solver with that macro and the reference
// headers define non-inline functions (signArray.h:19, histogramIO.h:230+),
// so both flavours cannot share one link unit.

#include "inverseWasserstein.h"

#include <cstring>
#include <type_traits>
#include <vector>

namespace {

// ExportHistogramBase (histogramIO.h:42-54) with every export switched off.
struct NullExport : public ExportHistogramBase {
  NullExport() : ExportHistogramBase(".") {}
  int exportHistogram(const std::vector<double>&, std::string) const { return 0; }
  int exportHistogramsBatch(const std::vector<std::vector<double> >&,
                            std::vector<std::string>) const { return 0; }
};

template <class KT>
BaseLoss<KT>* make_loss(int loss, KT& k) {
  // Loss indices follow main_dictionary_learning.cpp:189-197.
  switch (loss) {
    case 1: return new TVLoss<KT>(k);
    case 2: return new QuadraticLoss<KT>(k);
    case 3: return new WassersteinLoss<KT>(k);
    case 4: return new KLLoss<KT>(k);
    default: return new QuadraticLoss<KT>(k);
  }
}

std::vector<std::vector<double> > split_atoms(const double* atoms, int K, int N) {
  std::vector<std::vector<double> > b(K);
  for (int i = 0; i < K; ++i) b[i].assign(atoms + (size_t)i * N, atoms + (size_t)(i + 1) * N);
  return b;
}

template <class KT>
void run_solver(WassersteinRegression<KT>& reg, double* la, double* dic, double* bary, int id,
                std::false_type) {
  reg.compute_gradient_both(la, dic, bary, id);
}
template <class KT>
void run_solver(WassersteinRegression<KT>& reg, double* la, double* dic, double* bary, int id,
                std::true_type) {
  reg.compute_gradient_both_log_signArray(la, dic, bary, id);
}

// One compute_gradient_both* call for histogram `id` of `observed`.
template <class KT, bool LOGDOM>
void gradient_one(int W, int H, double gamma, int K, int L, int loss, const double* atoms,
                  const double* lambdas, const double* observed, int P, int id, const double* b0,
                  double* gradLa, double* gradDic, double* bary, double* bL) {
  const int N = W * H;
  KT kernel(gamma, W, H);
  std::vector<double> obs(observed, observed + (size_t)P * N);
  std::vector<std::vector<double> > bases = split_atoms(atoms, K, N);
  std::vector<double> w(lambdas, lambdas + K);
  Problem<KT> problem(bases, obs, kernel, &w[0], P);
  BaseLoss<KT>* l = make_loss<KT>(loss, problem.kernel);
  NullExport ex;
  WassersteinRegression<KT> reg(&problem, L, GRADIENT_SINKHORN, *l, ex, 1.0, false, false, true,
                                b0 != NULL);
  if (b0) {
    reg.b.resize((size_t)K * N * (L + 1));
    std::memcpy(&reg.b[0], b0, sizeof(double) * K * N);
  }
  run_solver(reg, gradLa, gradDic, bary, id, std::integral_constant<bool, LOGDOM>());
  if (bL) std::memcpy(bL, &reg.b[(size_t)K * N * L], sizeof(double) * K * N);
  delete l;
}

template <class KT>
double evaluate(int W, int H, double gamma, int K, int P, int L, int loss, double scaleDic,
                int exp_weight, const double* atoms_init, const double* observed,
                double* b_storage, double* b_temp, const double* x, double* g) {
  const int N = W * H;
  KT kernel(gamma, W, H);
  std::vector<double> obs(observed, observed + (size_t)P * N);
  std::vector<std::vector<double> > bases = split_atoms(atoms_init, K, N);
  std::vector<double> w((size_t)K, 1.0 / K);
  Problem<KT> problem(bases, obs, kernel, &w[0], P);
  BaseLoss<KT>* l = make_loss<KT>(loss, problem.kernel);
  NullExport ex;
  const bool warm = b_storage != NULL;
  WassersteinRegression<KT> reg(&problem, L, GRADIENT_SINKHORN, *l, ex, scaleDic, false, false,
                                true, warm);
  reg.exp_weight = exp_weight != 0;
  reg.iteration = 0;
  if (warm) {
    // The state regress_both sets up for a warm restart (inverseWasserstein.h:378-387).
    reg.b_storage.assign(b_storage, b_storage + (size_t)K * N * P);
    reg.b_temp.resize((size_t)K * N * P);
    reg.b.resize((size_t)K * N * (L + 1));
  }
  const int n = P * K + K * N;
  double f = WassersteinRegression<KT>::evaluate_both(&reg, x, g, n, 0.0);
  if (warm && b_temp) std::memcpy(b_temp, &reg.b_temp[0], sizeof(double) * K * N * P);
  delete l;
  return f;
}

}  // namespace

extern "C" {

#ifndef COMPUTE_BARYCENTER_LOG

// Gaussian2DKernel::kernel1d (kernels.h:632-641).
void ref_kernel_table(int W, int H, double gamma, double* out) {
  Gaussian2DKernel k(gamma, W, H);
  std::memcpy(out, k.kernel1d, sizeof(double) * std::max(W, H));
}

// Gaussian2DKernel::convolve (kernels.h:669-699).
void ref_convolve(int W, int H, double gamma, const double* u, double* out, int nvec) {
  Gaussian2DKernel k(gamma, W, H);
  k.convolve(u, out, nvec);
}

// Gaussian1DKernel::convolve (kernels.h:533-549).
void ref_convolve1d(int W, double gamma, const double* u, double* out, int nvec) {
  Gaussian1DKernel k(gamma, W);
  k.convolve(u, out, nvec);
}

// Gaussian3DKernel::convolve (kernels.h:762-807).
void ref_convolve3d(int W, int H, int D, double gamma, const double* u, double* out, int nvec) {
  Gaussian3DKernel k(gamma, W, H, D);
  k.convolve(u, out, nvec);
}

// GaussianSSE2DKernel::convolve (kernels.h:878-897, sse_helpers.h:403-423).
// Its dot product uses aligned AVX loads (sse_helpers.h:213), so rows are
// copied into 32-byte aligned storage, as the reference's own aligned_allocator
// buffers are (inverseWasserstein.h:550); rows stay aligned only when W % 4 == 0.
void ref_sse_convolve(int W, int H, double gamma, const double* u, double* out, int nvec) {
  GaussianSSE2DKernel k(gamma, W, H);
  const size_t n = (size_t)W * H * nvec;
  std::vector<double, aligned_allocator<double> > in(u, u + n), res(n);
  k.convolve(&in[0], &res[0], nvec);
  std::memcpy(out, &res[0], n * sizeof(double));
}

// LogSignArrayGaussian2DKernel::log_convolve (kernels.h:1073-1112).
void ref_log_convolve(int W, int H, double gamma, const double* u, double* out, int nvec) {
  LogSignArrayGaussian2DKernel k(gamma, W, H);
  k.log_convolve(u, out, nvec);
}

// LogSignArrayGaussian2DKernel::log_convolve_signArray (kernels.h:1117-1159).
void ref_log_convolve_signarray(int W, int H, double gamma, const double* u,
                                unsigned char* signs, double* out, int nvec) {
  LogSignArrayGaussian2DKernel k(gamma, W, H);
  k.log_convolve_signArray(u, signs, out, nvec);
}

// myAbsLog / myAbsExp (signArray.h:24-35) over an array.
void ref_abs_log(const double* v, unsigned char* signs, double* out, int n) {
  for (int i = 0; i < n; ++i) out[i] = myAbsLog(v[i], signs, (unsigned)i);
}
void ref_abs_exp(const double* v, unsigned char* signs, double* out, int n) {
  for (int i = 0; i < n; ++i) out[i] = myAbsExp(v[i], signs, (unsigned)i);
}

// WassersteinRegression<Gaussian2DKernel>::compute_gradient_both (inverseWasserstein.h:544-674).
void ref_gradient_both(int W, int H, double gamma, int K, int L, int loss, const double* atoms,
                       const double* lambdas, const double* observed, const double* b0,
                       double* gradLa, double* gradDic, double* bary, double* bL) {
  gradient_one<Gaussian2DKernel, false>(W, H, gamma, K, L, loss, atoms, lambdas, observed, 1, 0,
                                        b0, gradLa, gradDic, bary, bL);
}

// Same with the AVX kernel (the reference's "SSE" KernelType).
void ref_gradient_both_sse(int W, int H, double gamma, int K, int L, int loss, const double* atoms,
                           const double* lambdas, const double* observed, const double* b0,
                           double* gradLa, double* gradDic, double* bary, double* bL) {
  gradient_one<GaussianSSE2DKernel, false>(W, H, gamma, K, L, loss, atoms, lambdas, observed, 1,
                                           0, b0, gradLa, gradDic, bary, bL);
}

// compute_gradient_both_log_signArray with LogSignArrayGaussian2DKernel
// (inverseWasserstein.h:678-814).
void ref_gradient_both_log(int W, int H, double gamma, int K, int L, int loss,
                           const double* atoms, const double* lambdas, const double* observed,
                           const double* b0, double* gradLa, double* gradDic, double* bary,
                           double* bL) {
  gradient_one<LogSignArrayGaussian2DKernel, true>(W, H, gamma, K, L, loss, atoms, lambdas,
                                                   observed, 1, 0, b0, gradLa, gradDic, bary, bL);
}

// evaluate_both (inverseWasserstein.h:986-1145); kernel 0 = Gaussian2DKernel, 1 = SSE/AVX.
double ref_evaluate_both(int kernel, int W, int H, double gamma, int K, int P, int L, int loss,
                         double scaleDic, int exp_weight, const double* atoms_init,
                         const double* observed, double* b_storage, double* b_temp,
                         const double* x, double* g) {
  if (kernel == 1)
    return evaluate<GaussianSSE2DKernel>(W, H, gamma, K, P, L, loss, scaleDic, exp_weight,
                                         atoms_init, observed, b_storage, b_temp, x, g);
  return evaluate<Gaussian2DKernel>(W, H, gamma, K, P, L, loss, scaleDic, exp_weight, atoms_init,
                                    observed, b_storage, b_temp, x, g);
}

int ref_num_threads(void) {
#ifdef HAS_OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

// OpenMP threads of the reference's parallel loops (for the 1-thread baseline).
void ref_set_threads(int n) {
#ifdef HAS_OPENMP
  omp_set_num_threads(n);
#else
  (void)n;
#endif
}

#else  // COMPUTE_BARYCENTER_LOG

// evaluate_both built with COMPUTE_BARYCENTER_LOG (inverseWasserstein.h:1066-1068).
double ref_evaluate_both_log(int W, int H, double gamma, int K, int P, int L, int loss,
                             double scaleDic, int exp_weight, const double* atoms_init,
                             const double* observed, double* b_storage, double* b_temp,
                             const double* x, double* g) {
  return evaluate<LogSignArrayGaussian2DKernel>(W, H, gamma, K, P, L, loss, scaleDic, exp_weight,
                                                atoms_init, observed, b_storage, b_temp, x, g);
}

#endif

}  // extern "C"
