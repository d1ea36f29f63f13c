// Drop-in check (test-only): the reference's own templates, compiled from its
// unmodified headers, instantiated with the HIP KernelType of
// include/wdl/hip_kernel_type.hpp, against the same templates with the
// reference's CPU kernels; and the reference's lbfgs() driven once by the
// reference's evaluate_both and once by wdl_evaluate_both (the batched GPU
// callback).  Prints one JSON line of relative errors; exit status 0 when
// every error is within its tolerance.
//
// Built by oracle/Makefile into oracle/_ref/dropin_check; run by tests/ on a GPU.

#include "inverseWasserstein.h"

#include <cstdint>
#include <cstdio>
#include <random>
#include <sstream>
#include <string>
#include <type_traits>
#include <vector>

#include "wdl/hip_kernel_type.hpp"

namespace {

struct NullExport : public ExportHistogramBase {
  NullExport() : ExportHistogramBase(".") {}
  int exportHistogram(const std::vector<double>&, std::string) const { return 0; }
  int exportHistogramsBatch(const std::vector<std::vector<double> >&,
                            std::vector<std::string>) const { return 0; }
};

double uniform(std::mt19937_64& g) { return (double)(g() >> 11) * (1.0 / 9007199254740992.0); }

double rel(const std::vector<double>& a, const std::vector<double>& b) {
  double num = 0, den = 0;
  for (size_t i = 0; i < a.size(); ++i) {
    num += (a[i] - b[i]) * (a[i] - b[i]);
    den += b[i] * b[i];
  }
  return std::sqrt(num / (den > 0 ? den : 1));
}

struct Data {
  int W, H, K, P, L;
  double gamma;
  std::vector<std::vector<double> > atoms;
  std::vector<double> obs, lam;
};

Data make(int W, int H, int K, int P, int L, double gamma, uint64_t seed) {
  Data d{W, H, K, P, L, gamma, {}, {}, {}};
  std::mt19937_64 g(seed);
  const int N = W * H;
  d.obs.resize((size_t)P * N);
  for (int p = 0; p < P; ++p) {
    double s = 0;
    for (int j = 0; j < N; ++j) s += d.obs[(size_t)p * N + j] = 0.1 + uniform(g);
    for (int j = 0; j < N; ++j) d.obs[(size_t)p * N + j] /= s;
  }
  d.atoms.assign(K, std::vector<double>(N));
  for (int i = 0; i < K; ++i)
    for (int j = 0; j < N; ++j) d.atoms[i][j] = 0.05 + uniform(g);
  d.lam.resize((size_t)P * K);
  for (auto& v : d.lam) v = 0.1 + uniform(g);
  return d;
}

template <class KT>
void solve(WassersteinRegression<KT>& reg, double* gl, double* gd, double* q, std::false_type) {
  reg.compute_gradient_both(gl, gd, q, 0);
}
template <class KT>
void solve(WassersteinRegression<KT>& reg, double* gl, double* gd, double* q, std::true_type) {
  reg.compute_gradient_both_log_signArray(gl, gd, q, 0);
}

// compute_gradient_both (or its log form) for histogram 0, with the given
// kernel object (its N must be d.W * d.H).
template <class KT, bool LOGDOM>
void gradient(const Data& d, KT& kernel, std::vector<double>& gl, std::vector<double>& gd,
              std::vector<double>& bary) {
  const int N = d.W * d.H;
  std::vector<std::vector<double> > atoms = d.atoms;
  Problem<KT> problem(atoms, d.obs, kernel, const_cast<double*>(&d.lam[0]), d.P);
  QuadraticLoss<KT> loss;
  NullExport ex;
  WassersteinRegression<KT> reg(&problem, d.L, GRADIENT_SINKHORN, loss, ex, 1.0, false, false,
                                true, false);
  gl.assign(d.K, 0);
  gd.assign((size_t)d.K * N, 0);
  bary.assign(N, 0);
  solve(reg, &gl[0], &gd[0], &bary[0], std::integral_constant<bool, LOGDOM>());
}

template <class KT, bool LOGDOM>
void gradient(const Data& d, std::vector<double>& gl, std::vector<double>& gd,
              std::vector<double>& bary) {
  KT kernel(d.gamma, d.W, d.H);
  gradient<KT, LOGDOM>(d, kernel, gl, gd, bary);
}

std::vector<double> initial_x(const Data& d, double scaleDic) {
  // regress_both's transform of (weights, atoms) to L-BFGS variables (inverseWasserstein.h:389-401)
  const int N = d.W * d.H;
  std::vector<double> x((size_t)d.P * d.K + (size_t)d.K * N);
  for (int p = 0; p < d.P; ++p) {
    double s = 0;
    for (int i = 0; i < d.K; ++i) s += d.lam[(size_t)p * d.K + i];
    for (int i = 0; i < d.K; ++i) x[(size_t)p * d.K + i] = std::log(d.lam[(size_t)p * d.K + i] / s) * scaleDic;
  }
  for (int i = 0; i < d.K; ++i) {
    double s = 0;
    for (int j = 0; j < N; ++j) s += d.atoms[i][j];
    for (int j = 0; j < N; ++j) x[(size_t)d.P * d.K + (size_t)i * N + j] = std::log(d.atoms[i][j] / s);
  }
  return x;
}

lbfgs_parameter_t regress_params(int iters) {
  // The settings regress_both uses (inverseWasserstein.h:365-374).
  lbfgs_parameter_t prm;
  lbfgs_parameter_init(&prm);
  prm.linesearch = LBFGS_LINESEARCH_BACKTRACKING_STRONG_WOLFE;
  prm.epsilon = 1E-50;
  prm.max_linesearch = 20;
  prm.delta = 1E-50;
  prm.max_iterations = iters;
  return prm;
}

int quiet_progress(void*, const lbfgsfloatval_t*, const lbfgsfloatval_t*, const lbfgsfloatval_t,
                   const lbfgsfloatval_t, const lbfgsfloatval_t, const lbfgsfloatval_t, int, int,
                   int) {
  return 0;
}

}  // namespace

int main() {
  bool ok = true;
  char buf[512];
  std::string out = "{";
  // 1. compute_gradient_both: reference template with the HIP KernelType.
  {
    Data d = make(24, 20, 3, 2, 10, 1.0, 11);
    std::vector<double> gl1, gd1, q1, gl2, gd2, q2;
    gradient<Gaussian2DKernel, false>(d, gl1, gd1, q1);
    gradient<wdl::HipGaussian2DKernel, false>(d, gl2, gd2, q2);
    const double e1 = rel(gl2, gl1), e2 = rel(gd2, gd1), e3 = rel(q2, q1);
    ok = ok && e1 < 1e-5 && e2 < 1e-5 && e3 < 1e-5;
    std::snprintf(buf, sizeof buf, "\"gradient_both\": [%.3e, %.3e, %.3e], ", e1, e2, e3);
    out += buf;
  }
  // 2. compute_gradient_both_log_signArray with the HIP log KernelType.
  {
    Data d = make(16, 12, 3, 1, 6, 0.5, 12);
    std::vector<double> gl1, gd1, q1, gl2, gd2, q2;
    gradient<LogSignArrayGaussian2DKernel, true>(d, gl1, gd1, q1);
    gradient<wdl::HipLogSignArrayGaussian2DKernel, true>(d, gl2, gd2, q2);
    const double e1 = rel(gl2, gl1), e2 = rel(gd2, gd1), e3 = rel(q2, q1);
    ok = ok && e1 < 1e-10 && e2 < 1e-10 && e3 < 1e-10;
    std::snprintf(buf, sizeof buf, "\"gradient_both_log\": [%.3e, %.3e, %.3e], ", e1, e2, e3);
    out += buf;
  }
  // 4. compute_gradient_both on 1-D histograms (150 bins) with the HIP
  //    Gaussian1DKernel drop-in.
  {
    Data d = make(150, 1, 3, 1, 8, 4.0, 13);
    std::vector<double> gl1, gd1, q1, gl2, gd2, q2;
    Gaussian1DKernel k1(d.gamma, 150);
    wdl::HipGaussian1DKernel k2(d.gamma, 150);
    gradient<Gaussian1DKernel, false>(d, k1, gl1, gd1, q1);
    gradient<wdl::HipGaussian1DKernel, false>(d, k2, gl2, gd2, q2);
    const double e1 = rel(gl2, gl1), e2 = rel(gd2, gd1), e3 = rel(q2, q1);
    ok = ok && e1 < 1e-5 && e2 < 1e-5 && e3 < 1e-5;
    std::snprintf(buf, sizeof buf, "\"gradient_both_1d\": [%.3e, %.3e, %.3e], ", e1, e2, e3);
    out += buf;
  }
  // 5. ... and on 12 x 10 x 8 volumes with the HIP Gaussian3DKernel drop-in.
  {
    Data d = make(12, 10 * 8, 2, 1, 6, 2.0, 14);
    std::vector<double> gl1, gd1, q1, gl2, gd2, q2;
    Gaussian3DKernel k1(d.gamma, 12, 10, 8);
    wdl::HipGaussian3DKernel k2(d.gamma, 12, 10, 8);
    gradient<Gaussian3DKernel, false>(d, k1, gl1, gd1, q1);
    gradient<wdl::HipGaussian3DKernel, false>(d, k2, gl2, gd2, q2);
    const double e1 = rel(gl2, gl1), e2 = rel(gd2, gd1), e3 = rel(q2, q1);
    ok = ok && e1 < 1e-5 && e2 < 1e-5 && e3 < 1e-5;
    std::snprintf(buf, sizeof buf, "\"gradient_both_3d\": [%.3e, %.3e, %.3e], ", e1, e2, e3);
    out += buf;
  }
  // 3. The reference's lbfgs() with its own evaluate_both, and with the batched
  //    GPU callback wdl_evaluate_both, from the same start.
  {
    Data d = make(20, 20, 3, 4, 8, 1.0, 13);
    const int N = d.W * d.H, n = d.P * d.K + d.K * N;
    const double scaleDic = ((double)N / d.P) / 100.0;  // main_dictionary_learning.cpp:138
    const int iters = 4;
    std::vector<double> x1 = initial_x(d, scaleDic), x2 = x1;
    lbfgs_parameter_t prm = regress_params(iters);

    Gaussian2DKernel kernel(d.gamma, d.W, d.H);
    std::vector<std::vector<double> > atoms = d.atoms;
    Problem<Gaussian2DKernel> problem(atoms, d.obs, kernel, &d.lam[0], d.P);
    QuadraticLoss<Gaussian2DKernel> loss;
    NullExport ex;
    WassersteinRegression<Gaussian2DKernel> reg(&problem, d.L, GRADIENT_SINKHORN, loss, ex,
                                                scaleDic, false, false, true, false);
    lbfgsfloatval_t f1 = 0, f2 = 0;
    std::ostringstream sink;  // the reference's evaluate_both reports progress on std::cout
    std::streambuf* keep = std::cout.rdbuf(sink.rdbuf());
    const int r1 = lbfgs(n, &x1[0], &f1, WassersteinRegression<Gaussian2DKernel>::evaluate_both,
                         quiet_progress, &reg, &prm);
    std::cout.rdbuf(keep);

    wdl_solver* s = nullptr;
    wdl::detail::check(wdl_solver_create(wdl::detail::shared_context().get(), d.W, d.H, d.K, d.P,
                                         d.L, d.gamma, WDL_LOSS_QUADRATIC, 0, &s),
                       "wdl_solver_create");
    wdl::detail::check(wdl_solver_set_observed(s, &d.obs[0]), "wdl_solver_set_observed");
    wdl::detail::check(wdl_solver_set_scale(s, scaleDic, 1), "wdl_solver_set_scale");
    const int r2 = lbfgs(n, &x2[0], &f2, wdl_evaluate_both, quiet_progress, s, &prm);
    wdl_solver_destroy(s);
    const double ex_ = rel(x2, x1), ef = std::fabs(f2 - f1) / std::fabs(f1);
    ok = ok && ef < 1e-4 && ex_ < 1e-4 && (r1 >= 0) == (r2 >= 0);
    std::snprintf(buf, sizeof buf,
                  "\"lbfgs\": {\"iters\": %d, \"ret_ref\": %d, \"ret_gpu\": %d, \"f_ref\": %.10e, "
                  "\"f_gpu\": %.10e, \"rel_f\": %.3e, \"rel_x\": %.3e}, ",
                  iters, r1, r2, f1, f2, ef, ex_);
    out += buf;
  }
  out += ok ? "\"ok\": true}" : "\"ok\": false}";
  std::cout << std::flush;
  std::printf("\nDROPIN %s\n", out.c_str());
  return ok ? 0 : 1;
}
