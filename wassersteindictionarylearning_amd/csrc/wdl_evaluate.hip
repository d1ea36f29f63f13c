// The evaluate_both epilogue and prologue on the device
// (WassersteinRegression::evaluate_both, inverseWasserstein.h:986-1145):
// x -> simplex weights and atoms, and after the batched solve the loss sum,
// the simplex-tangent projections and the chain rule of the reparametrisation,
// so x and g (P*K + K*N doubles) are the only data an L-BFGS evaluation moves.
// Everything here is FP64, as in the reference.
#include <cmath>

#include "wdl_solve_common.h"

namespace wdl {

// Weights (:998-1012): v = exp(x / scale) (exp_weight) or x / scale, then
// normalised per histogram.  w64 keeps that vector for the projections; the
// solve's FP32 copy is normalised once more, as normalize_values does again
// at the start of every solve (:53-74, :562).
__global__ void k_ev_weights(const double* __restrict__ x, double* __restrict__ w64,
                             float* __restrict__ lam, int P, int K, double scale, int expw) {
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < P; p += gridDim.x * blockDim.x) {
    const double* xp = x + (long long)p * K;
    double* wp = w64 + (long long)p * K;
    double t = 0;
    for (int i = 0; i < K; ++i) {
      const double v = expw ? exp(xp[i] / scale) : xp[i] / scale;
      wp[i] = v;
      t += v;
    }
    double t2 = 0;
    for (int i = 0; i < K; ++i) {
      wp[i] /= t;
      t2 += wp[i];
    }
    for (int i = 0; i < K; ++i) lam[(long long)p * K + i] = (float)(wp[i] / t2);
  }
}

// Atoms (:1013-1031): a = exp(x) normalised per atom; the solve's FP32 copy
// (and its log for the log domain, :704) is normalised once more as above.
// Each atom is split into kEvChunk-pixel chunks, one workgroup per (atom,
// chunk), so the K atoms spread over many CUs; a per-atom total is the sum of
// its chunks' partials in chunk order (deterministic), re-added by every
// workgroup of the atom in the next launch.  part: 2 x K x chunks doubles.
__device__ inline void ev_chunk(long long N, long long& base, long long& j0, long long& j1,
                                int& i, int& c) {
  const int nc = (int)ev_chunks(N);
  i = blockIdx.x / nc;
  c = blockIdx.x - i * nc;
  base = (long long)i * N;
  j0 = (long long)c * kEvChunk;
  j1 = j0 + kEvChunk < N ? j0 + kEvChunk : N;
}

__device__ inline double ev_total(const double* __restrict__ part, int i, long long N) {
  __shared__ double tot;
  const int nc = (int)ev_chunks(N);
  if (threadIdx.x == 0) {
    double t = 0;
    for (int c = 0; c < nc; ++c) t += part[(long long)i * nc + c];
    tot = t;
  }
  __syncthreads();
  return tot;
}

__global__ __launch_bounds__(kThreads) void k_ev_atoms_exp(const double* __restrict__ xa,
                                                           double* __restrict__ a64,
                                                           double* __restrict__ part, long long N) {
  long long base, j0, j1;
  int i, c;
  ev_chunk(N, base, j0, j1, i, c);
  double t = 0;
  for (long long j = j0 + threadIdx.x; j < j1; j += kThreads) {
    const double e = exp(xa[base + j]);
    a64[base + j] = e;
    t += e;
  }
  t = block_sum(t);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

__global__ __launch_bounds__(kThreads) void k_ev_atoms_norm(double* __restrict__ a64,
                                                            double* __restrict__ part,
                                                            long long N) {
  long long base, j0, j1;
  int i, c;
  ev_chunk(N, base, j0, j1, i, c);
  const double tot = ev_total(part, i, N);
  double t2 = 0;
  for (long long j = j0 + threadIdx.x; j < j1; j += kThreads) {
    const double v = a64[base + j] / tot;
    a64[base + j] = v;
    t2 += v;
  }
  t2 = block_sum(t2);
  if (threadIdx.x == 0) part[gridDim.x + blockIdx.x] = t2;
}

__global__ __launch_bounds__(kThreads) void k_ev_atoms_f32(const double* __restrict__ a64,
                                                           const double* __restrict__ part,
                                                           float* __restrict__ A,
                                                           float* __restrict__ logA, long long N) {
  long long base, j0, j1;
  int i, c;
  ev_chunk(N, base, j0, j1, i, c);
  const double tot2 = ev_total(part + gridDim.x, i, N);
  for (long long j = j0 + threadIdx.x; j < j1; j += kThreads) {
    const double v = a64[base + j] / tot2;
    A[base + j] = (float)v;
    if (logA) logA[base + j] = (float)log(v);
  }
}

// f = sum of the per-histogram losses (:1089-1091); weight gradients projected
// on the simplex tangent, g = G - <w, G> (:1095-1101), times w under
// exp_weight, over scale (:1125-1136).  One workgroup.
__global__ __launch_bounds__(kThreads) void k_ev_weights_grad(
    const double* __restrict__ GL, const double* __restrict__ w64,
    const double* __restrict__ lossPart, double* __restrict__ g, double* __restrict__ f, int P,
    int K, int nLoss, double scale, int expw) {
  double t = 0;
  for (long long e = threadIdx.x; e < (long long)P * nLoss; e += kThreads) t += lossPart[e];
  t = block_sum(t);
  if (threadIdx.x == 0) *f = t;
  for (int p = threadIdx.x; p < P; p += kThreads) {
    const long long o = (long long)p * K;
    double d = 0;
    for (int i = 0; i < K; ++i) d += w64[o + i] * GL[o + i];
    for (int i = 0; i < K; ++i) {
      double v = GL[o + i] - d;
      if (expw) v *= w64[o + i];
      g[o + i] = v / scale;
    }
  }
}

// Atom gradients (:1103-1112): the projection G - <a_i, G> is linear in the
// per-histogram gradients, so it is applied to their sum; times a for the
// softmax (:1125-1136).  Chunked as the atoms above: k_ev_atoms_dot writes
// the chunks' partial dots, k_ev_atoms_grad adds them in order.  `zero`:
// weights-only solve, whose atom gradients are zero.
__global__ __launch_bounds__(kThreads) void k_ev_atoms_dot(const double* __restrict__ GDsum,
                                                           const double* __restrict__ a64,
                                                           double* __restrict__ part, long long N) {
  long long base, j0, j1;
  int i, c;
  ev_chunk(N, base, j0, j1, i, c);
  double d = 0;
  for (long long j = j0 + threadIdx.x; j < j1; j += kThreads) d += a64[base + j] * GDsum[base + j];
  d = block_sum(d);
  if (threadIdx.x == 0) part[blockIdx.x] = d;
}

__global__ __launch_bounds__(kThreads) void k_ev_atoms_grad(const double* __restrict__ GDsum,
                                                            const double* __restrict__ a64,
                                                            const double* __restrict__ part,
                                                            double* __restrict__ g, long long N,
                                                            int zero) {
  long long base, j0, j1;
  int i, c;
  ev_chunk(N, base, j0, j1, i, c);
  const double d = zero ? 0.0 : ev_total(part, i, N);
  for (long long j = j0 + threadIdx.x; j < j1; j += kThreads)
    g[base + j] = zero ? 0.0 : (GDsum[base + j] - d) * a64[base + j];
}

int evaluate_prologue(const double* x, double* w64, double* a64, float* lam, float* A,
                      float* logA, double* part, int P, int K, long long N, double scale,
                      int expw, hipStream_t st) {
  const dim3 ga((unsigned)(K * ev_chunks(N)));
  hipLaunchKernelGGL(k_ev_weights, dim3((P + 255) / 256), dim3(256), 0, st, x, w64, lam, P, K,
                     scale, expw);
  const double* xa = x + (long long)P * K;
  hipLaunchKernelGGL(k_ev_atoms_exp, ga, dim3(kThreads), 0, st, xa, a64, part, N);
  hipLaunchKernelGGL(k_ev_atoms_norm, ga, dim3(kThreads), 0, st, a64, part, N);
  hipLaunchKernelGGL(k_ev_atoms_f32, ga, dim3(kThreads), 0, st, a64, part, A, logA, N);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int evaluate_epilogue(const double* GL, const double* GDsum, const double* lossPart,
                      const double* w64, const double* a64, double* part, double* g, double* f,
                      int P, int K, long long N, double scale, int expw, int weightsOnly,
                      hipStream_t st) {
  const dim3 ga((unsigned)(K * ev_chunks(N)));
  hipLaunchKernelGGL(k_ev_weights_grad, dim3(1), dim3(kThreads), 0, st, GL, w64, lossPart, g, f,
                     P, K, kLossBlocks, scale, expw);
  if (!weightsOnly)
    hipLaunchKernelGGL(k_ev_atoms_dot, ga, dim3(kThreads), 0, st, GDsum, a64, part, N);
  hipLaunchKernelGGL(k_ev_atoms_grad, ga, dim3(kThreads), 0, st, GDsum, a64, part,
                     g + (long long)P * K, N, weightsOnly);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace wdl
