// Exact-shift log-domain convolution, and the log-domain solver stages built
// on it (WDL_FLAG_EXACT_LSE).
//
// Each 1-D pass computes, for every output j of a line x,
//     m_j = max(-1e99, max_k(x_k - (j-k)^2/gamma))               (kernels.h:1082-1091)
//     out_j = m_j + log |sum_k s_k exp(x_k - (j-k)^2/gamma - m_j)|
// with the reference's exact per-output maximum, so any spread of log values
// along a line is handled (the shifted MFMA form of wdl_logconv.h needs every
// line within ~80 nats).  The line is split into 16-pixel chunks with their
// maxima; a chunk whose bound max_c - d_min^2/gamma cannot raise m_j is not
// visited for the maximum, and one that lies more than 40 nats below m_j is
// not visited for the sum (its terms are below e^-40 each: at most len * 4e-18
// relative).  Terms are formed in FP64, exponentiated and summed in FP32 (each
// term <= ~1).  Cost per output: one bound per chunk plus the visited chunks'
// terms -- O(len) per line in the worst case, a few chunks when the log values
// are smooth.
//
// The solver stages (exact_forward / exact_backward) run the log-domain
// recursions of compute_gradient_both_log_signArray (inverseWasserstein.h:
// 678-814) as elementwise kernels around these passes, with FP64 scratch
// images between them.
#include <algorithm>
#include <cmath>

#include "wdl_solve_common.h"

namespace wdl {

constexpr int kXWaves = 4;  // lines per workgroup, one wave each
constexpr double kXPrune = 40.0;

__host__ __device__ inline int xlse_lds_bytes(int len) {
  const int nch = (len + 15) / 16;
  return kXWaves * (len * (8 + 8 + 4 + 4 + 1) + nch * (8 + 4)) + 64;
}

// Reach, in chunks, of the 16 outputs of a chunk: with m the smallest output
// maximum among them (lanes with `valid` false excluded), a chunk c with
// |c - oc| > R has bound <= gmax - (16 (|c - oc| - 1))^2 / gamma < m - margin.
__device__ inline int chunk_reach(float gmax, float m, bool valid, float margin, float gam,
                                  int nch) {
  float mm = valid ? m : INFINITY;
  for (int sh = 1; sh < 16; sh <<= 1) mm = fminf(mm, __shfl_xor(mm, sh));
  const float span = gmax - mm + margin;
  if (!(span < 1e30f)) return nch;
  if (span <= 0.f) return 1;
  return min(nch, 1 + (int)ceilf(sqrtf(span * gam) * (1.f / 16.f)));
}

// One pass over `nlines` lines of length `len` (line-major input), writing the
// transposed layout (line index fastest) of each image of `rows` lines.  A line
// is skipped (output untouched) when active_thr > 0 and Lp[p0 + image / ipp] <
// active_thr.
//
// One wave per line.  Outputs are taken 16 at a time (one output chunk): lane l
// serves output j = 16 oc + (l % 16) over the quarter l / 16 of every visited
// input chunk, and the four lanes of an output combine their maxima and sums
// with cross-lane shuffles, so the 16 outputs of a chunk -- which need nearly
// the same input chunks -- move through them together.  The maximum is found
// in FP32 (any shift near the maximum serves: the output adds back exactly the
// shift the terms were formed with); terms are formed in FP64.
template <bool SIGNED>
__global__ __launch_bounds__(kXWaves * 64) void k_xlse_pass(
    const double* __restrict__ in, const unsigned char* __restrict__ sin, double* __restrict__ out,
    unsigned char* __restrict__ sout, int rows, int len, long long nlines, double inv_gamma,
    const int* __restrict__ Lp, int p0, int ipp, int active_thr) {
  extern __shared__ __align__(16) double xs[];
  const int nch = (len + 15) >> 4;
  const int wave = wave_id(), lane = threadIdx.x & 63;
  const int o = lane & 15, q4 = lane >> 4;
  const float igf = (float)inv_gamma;
  const float gam = (float)(1.0 / inv_gamma);
  // per wave: its line X (FP64) and XF (FP32), chunk maxima CM / CMF, input
  // signs SG, outputs O and their signs SO; O0 / SO0 address all waves' outputs
  double* X = xs + wave * len;
  double* O0 = xs + kXWaves * len;
  double* O = O0 + wave * len;
  double* CM = xs + 2 * kXWaves * len + wave * nch;
  float* F0 = reinterpret_cast<float*>(xs + 2 * kXWaves * len + kXWaves * nch);
  float* SG = F0 + wave * len;
  float* XF = F0 + kXWaves * len + wave * len;
  float* CMF = F0 + 2 * kXWaves * len + wave * nch;
  unsigned char* SO0 = reinterpret_cast<unsigned char*>(F0 + 2 * kXWaves * len + kXWaves * nch);
  unsigned char* SO = SO0 + wave * len;
  for (long long l0 = (long long)blockIdx.x * kXWaves; l0 < nlines;
       l0 += (long long)gridDim.x * kXWaves) {
    const long long line = l0 + wave;
    bool active = line < nlines;
    if (active && active_thr > 0) {
      const long long img = line / rows;
      active = Lp[p0 + (int)(img / ipp)] >= active_thr;
    }
    float gmax = -INFINITY;  // the line's maximum, in every lane
    if (active) {
      // 64 values per step; each 16-lane group holds one chunk and reduces
      // its maximum with shuffles (all lanes take part in every step)
      const double* src = in + line * len;
      for (int base = 0; base < len; base += 64) {
        const int k = base + lane;
        const bool in_line = k < len;
        const double x = in_line ? src[k] : -INFINITY;
        if (in_line) {
          X[k] = x;
          XF[k] = (float)x;
          if (SIGNED) SG[k] = sin[line * len + k] ? 1.f : -1.f;  // getSign (signArray.h:19-22)
        }
        double m = x;
        for (int sh = 1; sh < 16; sh <<= 1) m = fmax(m, __shfl_xor(m, sh));
        if ((lane & 15) == 0 && k < len) {
          CM[k >> 4] = m;
          CMF[k >> 4] = (float)m;
        }
        gmax = fmaxf(gmax, (float)m);
      }
      gmax = fmaxf(gmax, __shfl_xor(gmax, 16));
      gmax = fmaxf(gmax, __shfl_xor(gmax, 32));
    }
    __syncthreads();
    if (active)
      for (int oc = 0; oc < nch; ++oc) {
        const int j = oc * 16 + o;
        // The lane's quarter of chunk c is k = kb..kb+3, kb = 16c + 4q4; a
        // whole quarter reads its four values together and needs no guard.
        auto qmax = [&](int c, float r) {
          const int kb = c * 16 + q4 * 4;
          if (kb + 4 <= len) {
            float x[4];
#pragma unroll
            for (int t = 0; t < 4; ++t) x[t] = XF[kb + t];
#pragma unroll
            for (int t = 0; t < 4; ++t) {
              const int dk = j - kb - t;
              r = fmaxf(r, x[t] - (float)__mul24(dk, dk) * igf);
            }
          } else {
            for (int t = 0; t < 4; ++t) {
              const int k = kb + t;
              if (k < len) r = fmaxf(r, XF[k] - (float)((j - k) * (j - k)) * igf);
            }
          }
          return r;
        };
        // maximum of x_k - (j-k)^2/gamma: own chunk, then every chunk whose
        // bound can raise it (the four lanes of j decide alike)
        float m = qmax(oc, -INFINITY);
        m = fmaxf(m, __shfl_xor(m, 16));
        m = fmaxf(m, __shfl_xor(m, 32));
        // chunks more than R away cannot raise the maximum of any of the 16
        // outputs: their bound is at most gmax - (16 (R-1))^2 / gamma <= m
        const int R = chunk_reach(gmax, m, j < len, 0.f, gam, nch);
        for (int c = max(0, oc - R); c <= min(nch - 1, oc + R); ++c) {
          if (c == oc) continue;
          const int d = c < oc ? j - (c * 16 + 15) : c * 16 - j;
          if (CMF[c] - (float)(d * d) * igf <= m) continue;
          float mc = qmax(c, -INFINITY);
          mc = fmaxf(mc, __shfl_xor(mc, 16));
          mc = fmaxf(mc, __shfl_xor(mc, 32));
          m = fmaxf(m, mc);
        }
        const double m64 = fmax((double)m, -1e99);
        float sum = 0.f;
        // chunks beyond Rs lie more than 40 nats below every output's maximum
        const int Rs = chunk_reach(gmax, m, j < len, (float)kXPrune + 1.f, gam, nch);
        for (int c = max(0, oc - Rs); c <= min(nch - 1, oc + Rs); ++c) {
          const int d = c == oc ? 0 : (c < oc ? j - (c * 16 + 15) : c * 16 - j);
          if (CM[c] - (double)(d * d) * inv_gamma - m64 < -kXPrune) continue;
          // the term's exponent is formed in FP64 (exact however large X and m
          // are), then exp(x) = 2^(x log2 e) on the hardware exp2: terms that
          // matter have x in [-40, 0], where the FP32 product rounds by at most
          // ~3.5e-6 relative (~3e-7 on the terms that carry the sum)
          auto term = [&](double xk, int k) {
            const int dk = j - k;
            const float x = (float)(xk - (double)__mul24(dk, dk) * inv_gamma - m64);
            return __builtin_amdgcn_exp2f(x * 1.4426950408889634f);
          };
          const int kb = c * 16 + q4 * 4;
          if (kb + 4 <= len) {
            double xk[4];
            float sg[4];
#pragma unroll
            for (int t = 0; t < 4; ++t) {
              xk[t] = X[kb + t];
              if (SIGNED) sg[t] = SG[kb + t];
            }
#pragma unroll
            for (int t = 0; t < 4; ++t) {
              const float e = term(xk[t], kb + t);
              sum += SIGNED ? sg[t] * e : e;
            }
          } else {
            for (int t = 0; t < 4; ++t) {
              const int k = kb + t;
              if (k < len) {
                const float e = term(X[k], k);
                sum += SIGNED ? SG[k] * e : e;
              }
            }
          }
        }
        sum += __shfl_xor(sum, 16);
        sum += __shfl_xor(sum, 32);
        if (q4 == 0 && j < len) {
          O[j] = m64 + hlog(fabsf(sum));
          if (SIGNED) SO[j] = sum >= 0.f ? 1 : 0;
        }
      }
    __syncthreads();
    // transposed store: the lines of one workgroup are consecutive, so each
    // output position j gets up to kXWaves adjacent values
    for (int e = threadIdx.x; e < kXWaves * len; e += blockDim.x) {
      const int w = e % kXWaves, j = e / kXWaves;
      const long long ln = l0 + w;
      if (ln >= nlines) continue;
      const long long img = ln / rows;
      if (active_thr > 0 && Lp[p0 + (int)(img / ipp)] < active_thr) continue;
      const int r = (int)(ln - img * rows);
      const long long oi = img * (long long)rows * len + (long long)j * rows + r;
      out[oi] = O0[w * len + j];
      if (SIGNED) sout[oi] = SO0[w * len + j];
    }
    __syncthreads();
  }
}

// Two passes (along x, then along y) over nimg W x H images; tmp holds the
// transposed intermediate.  Signs: none (sin == nullptr) or one byte per value.
static int xlse_conv(const double* in, const unsigned char* sin, double* out, unsigned char* sout,
                     double* tmp, unsigned char* stmp, int W, int H, long long nimg, double gamma,
                     const int* Lp, int p0, int ipp, int thr, hipStream_t st) {
  const double ig = 1.0 / gamma;
  const int thr_ = Lp ? thr : 0;
  auto launch = [&](const double* i, const unsigned char* si, double* o, unsigned char* so,
                    int rows, int len) {
    const long long nl = nimg * rows;
    const int grid = (int)std::min<long long>((nl + kXWaves - 1) / kXWaves, 1 << 20);
    const int lds = xlse_lds_bytes(len);
    if (si)
      hipLaunchKernelGGL(k_xlse_pass<true>, dim3(grid), dim3(kXWaves * 64), lds, st, i, si, o, so,
                         rows, len, nl, ig, Lp, p0, ipp, thr_);
    else
      hipLaunchKernelGGL(k_xlse_pass<false>, dim3(grid), dim3(kXWaves * 64), lds, st, i, si, o,
                         so, rows, len, nl, ig, Lp, p0, ipp, thr_);
  };
  launch(in, sin, tmp, sin ? stmp : nullptr, H, W);   // along x: lines are rows
  launch(tmp, sin ? stmp : nullptr, out, sout, W, H);  // along y: lines are columns
  WDL_HIP(hipGetLastError());
  return WDL_OK;
}

int xlse_set_lds_attr(int maxlen) {
  static int done = 0;
  if (done >= maxlen) return WDL_OK;
  const int b = xlse_lds_bytes(maxlen);
  WDL_HIP(hipFuncSetAttribute((const void*)k_xlse_pass<true>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, b));
  WDL_HIP(hipFuncSetAttribute((const void*)k_xlse_pass<false>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, b));
  done = maxlen;
  return WDL_OK;
}

int xlse_convolve(const double* in, const unsigned char* sin, double* out, unsigned char* sout,
                  double* tmp, unsigned char* stmp, int W, int H, long long nimg, double gamma,
                  hipStream_t st) {
  int e = xlse_set_lds_attr(std::max(W, H));
  if (e) return e;
  return xlse_conv(in, sin, out, sout, tmp, stmp, W, H, nimg, gamma, nullptr, 0, 1, 0, st);
}

// ---- solver stages -------------------------------------------------------------
// Scratch layout: image ((p * nch + chain slot) * K + i) of N values, nch = 2
// (atom chain, weight chain) or 1 (forward; weights-only backward), so the
// images of histograms p0..p0+np-1 are contiguous.

struct XEl {
  int p, i, cs;
  long long x, img;  // pixel, (p, i) image offset into the P x K x N arrays
};

__device__ inline XEl xel(const SolveArgs& a, int nch, long long e) {
  XEl r;
  const long long N = a.N;
  r.x = e % N;
  const long long im = e / N;
  r.i = (int)(im % a.K);
  const long long pc = im / a.K;
  r.cs = (int)(pc % nch);
  r.p = a.p0 + (int)(pc / nch);
  r.img = ((long long)r.p * a.K + r.i) * N;
  return r;
}

__device__ inline long long xoff(const SolveArgs& a, int nch, const XEl& r) {
  return (((long long)r.p * nch + r.cs) * a.K + r.i) * a.N + r.x;
}

#define XLOOP(total)                                                                     \
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < (total);      \
       e += (long long)gridDim.x * blockDim.x)

// forward, input of the first convolution: log b^{s-1} (:711-724; b^0: 0 or the store)
__global__ void k_xf1(SolveArgs a, int s) {
  const long long tot = (long long)a.np * a.K * a.N;
  XLOOP(tot) {
    const XEl r = xel(a, 1, e);
    double b;
    if (s > steps_of(a, r.p)) continue;
    if (s == 1) {
      b = a.bInit ? (double)a.bInit[r.img + r.x] : 0.0;
    } else {
      const double* Fp = (const double*)a.Fh + (long long)(s - 2) * a.stage +
                         (long long)r.p * a.K * a.N + r.x;
      double lq = 0;
      const float* lam = a.lam + (long long)r.p * a.K;
      for_atoms_l<4>(
          a.K, [&](int j) { return Fp[j * a.N]; },
          [&](int j, double f) { lq += (double)lam[j] * f; });
      b = lq - Fp[(long long)r.i * a.N];
      ((double*)a.Bh)[(long long)(s - 2) * a.stage + r.img + r.x] = b;
    }
    a.xa[xoff(a, 1, r)] = b;
  }
}

// forward, input of the second convolution: log p - log c^{s-1} (:704-707)
__global__ void k_xf2(SolveArgs a, int s) {
  const long long tot = (long long)a.np * a.K * a.N;
  XLOOP(tot) {
    const XEl r = xel(a, 1, e);
    if (s > steps_of(a, r.p)) continue;
    const double c = ((const double*)a.Ch)[(long long)(s - 1) * a.stage + r.img + r.x];
    a.xa[xoff(a, 1, r)] = (double)a.logA[(long long)r.i * a.N + r.x] - c;
  }
}

// backward step, first convolution's input: chain 0 (lam_i n + v_i) b^s, chain 1
// (lam_i g + r_i) / phi^s, as log magnitude and sign (:740-749, :778-793)
__global__ void k_xb1(SolveArgs a, int t, int nch) {
  const long long tot = (long long)a.np * nch * a.K * a.N;
  XLOOP(tot) {
    XEl r = xel(a, nch, e);
    const int s = bwd_step(a, r.p, t);
    if (s < 1) continue;
    const int chain = nch == 1 ? 1 : r.cs;
    const bool last = s == steps_of(a, r.p);
    const long long N = a.N, hb = (long long)r.p * a.K * N;
    const float lpi = a.lam[(long long)r.p * a.K + r.i];
    float acc = 0.f, own = 0.f;
    const float* Pp = chain == 0 ? ((s & 1) ? a.V0 : a.V1) : ((s & 1) ? a.R0 : a.R1);
    if (last) {
      acc = a.G0[(long long)r.p * N + r.x];
      if (chain == 1) acc *= a.Q[(long long)r.p * N + r.x];
    } else {
      for_atoms_l(
          a.K, [&](int j) { return Pp[hb + j * N + r.x]; },
          [&](int j, float v) {
            acc -= v;  // n = -sum v (:764), g = -sum r (:808)
            if (j == r.i) own = v;
          });
    }
    const float xv = lpi * acc + own;
    const long long h = (long long)(s - 1) * a.stage + r.img + r.x;
    const double lh = chain == 0 ? ((const double*)a.Bh)[h] : -((const double*)a.Fh)[h];
    const long long o = xoff(a, nch, r);
    a.xa[o] = hlog(fabsf(xv)) + lh;
    a.sga[o] = xv >= 0.f ? 1 : 0;
  }
}

// backward step, after the first convolution: resultDic (chain 0, :750-752) and
// the second convolution's input log p + c - 2 log c^{s-1}, sign of c (:754-756, :796)
__global__ void k_xb2(SolveArgs a, int t, int nch) {
  const long long tot = (long long)a.np * nch * a.K * a.N;
  XLOOP(tot) {
    XEl r = xel(a, nch, e);
    const int s = bwd_step(a, r.p, t);
    if (s < 1) continue;
    const int chain = nch == 1 ? 1 : r.cs;
    const long long o = xoff(a, nch, r);
    const double c = a.xb[o];
    const float sg = a.sgb[o] ? 1.f : -1.f;
    const double cs = ((const double*)a.Ch)[(long long)(s - 1) * a.stage + r.img + r.x];
    if (chain == 0) {
      float* gd = a.GD + r.img + r.x;
      const float prev = s == steps_of(a, r.p) ? 0.f : *gd;
      *gd = prev + sg * expf((float)(c - cs));
    }
    if (s > 1) {
      a.xa[o] = (double)a.logA[(long long)r.i * a.N + r.x] + c - 2.0 * cs;
      a.sga[o] = a.sgb[o];
    }
  }
}

// backward step s >= 2, after the second convolution: -v = sign exp(res - log
// phi^{s-1}) (:757-762), -r = sign exp(res + log b^{s-1}) (:797-805)
__global__ void k_xb3(SolveArgs a, int t, int nch) {
  const long long tot = (long long)a.np * nch * a.K * a.N;
  XLOOP(tot) {
    XEl r = xel(a, nch, e);
    const int s = bwd_step(a, r.p, t);
    if (s < 2) continue;
    const int chain = nch == 1 ? 1 : r.cs;
    const long long o = xoff(a, nch, r);
    const double v = a.xb[o];
    const float sg = a.sgb[o] ? 1.f : -1.f;
    if (chain == 0) {
      const double f = ((const double*)a.Fh)[(long long)(s - 2) * a.stage + r.img + r.x];
      ((s & 1) ? a.V1 : a.V0)[r.img + r.x] = sg * expf((float)(v - f));
    } else {
      const double b = ((const double*)a.Bh)[(long long)(s - 2) * a.stage + r.img + r.x];
      ((s & 1) ? a.R1 : a.R0)[r.img + r.x] = sg * expf((float)(v + b));
    }
  }
}

// weight-gradient dot of backward step s, <log phi_i^s, g_s> (:786), centred as
// in the other solvers (exact_gsum); one workgroup per (histogram, atom)
__global__ __launch_bounds__(kThreads) void k_xdot(SolveArgs a, int t) {
  const int p = a.p0 + (int)(blockIdx.x / a.K), i = (int)(blockIdx.x % a.K);
  const int s = bwd_step(a, p, t);
  if (s < 1) return;
  const bool last = s == steps_of(a, p);
  const long long N = a.N, hb = (long long)p * a.K * N;
  const double* F = (const double*)a.Fh + (long long)(s - 1) * a.stage + hb + (long long)i * N;
  const float* Rp = (s & 1) ? a.R0 : a.R1;
  double sfg = 0, sf = 0, sg = 0;
  for (long long x = threadIdx.x; x < N; x += blockDim.x) {
    float g = 0.f;
    if (last) {
      g = a.G0[(long long)p * N + x] * a.Q[(long long)p * N + x];
    } else {
      for_atoms_l(a.K, [&](int j) { return Rp[hb + j * N + x]; }, [&](int, float v) { g -= v; });
    }
    sfg += F[x] * (double)g;
    sf += F[x];
    sg += (double)g;
  }
  const double tfg = block_sum(sfg), tf = block_sum(sf), tg = block_sum(sg);
  if (threadIdx.x == 0) {
    double* gl = a.GL + (long long)p * a.K + i;
    *gl = last ? tfg : *gl + (tfg - tf / (double)N * (tg - exact_gsum(a, p, s)));
  }
}

static dim3 ew_grid(long long n) {
  return dim3((unsigned)std::min<long long>((n + 255) / 256, 8192));
}

int exact_forward(const SolveArgs& a, hipStream_t st, int s) {
  const long long nimg = (long long)a.np * a.K, n = nimg * a.N;
  const long long off = (long long)a.p0 * a.K * a.N;
  const int W = a.G.W, H = a.G.H;
  const double g = a.gamma;
  int e;
  hipLaunchKernelGGL(k_xf1, ew_grid(n), dim3(256), 0, st, a, s);
  if ((e = xlse_conv(a.xa + off, nullptr, (double*)a.Ch + (long long)(s - 1) * a.stage + off,
                     nullptr, a.xt + off, nullptr, W, H, nimg, g, a.Lp, a.p0, a.K, s, st)))
    return e;
  hipLaunchKernelGGL(k_xf2, ew_grid(n), dim3(256), 0, st, a, s);
  if ((e = xlse_conv(a.xa + off, nullptr, (double*)a.Fh + (long long)(s - 1) * a.stage + off,
                     nullptr, a.xt + off, nullptr, W, H, nimg, g, a.Lp, a.p0, a.K, s, st)))
    return e;
  WDL_HIP(hipGetLastError());
  return WDL_OK;
}

int exact_backward(const SolveArgs& a, hipStream_t st, int t) {
  const int nch = a.weightsOnly ? 1 : 2;
  const long long nimg = (long long)a.np * nch * a.K, n = nimg * a.N;
  const long long off = (long long)a.p0 * nch * a.K * a.N;
  const int W = a.G.W, H = a.G.H;
  const double g = a.gamma;
  int e;
  hipLaunchKernelGGL(k_xdot, dim3((unsigned)(a.np * a.K)), dim3(kThreads), 0, st, a, t);
  hipLaunchKernelGGL(k_xb1, ew_grid(n), dim3(256), 0, st, a, t, nch);
  // histogram p runs this step when L_p >= L - t + 1, the second convolution when >= L - t + 2
  if ((e = xlse_conv(a.xa + off, a.sga + off, a.xb + off, a.sgb + off, a.xt + off, a.sgt + off, W, H,
                     nimg, g, a.Lp, a.p0, nch * a.K, a.L - t + 1, st)))
    return e;
  hipLaunchKernelGGL(k_xb2, ew_grid(n), dim3(256), 0, st, a, t, nch);
  if (t >= 2) {
    if ((e = xlse_conv(a.xa + off, a.sga + off, a.xb + off, a.sgb + off, a.xt + off, a.sgt + off, W,
                       H, nimg, g, a.Lp, a.p0, nch * a.K, a.L - t + 2, st)))
      return e;
    hipLaunchKernelGGL(k_xb3, ew_grid(n), dim3(256), 0, st, a, t, nch);
  }
  WDL_HIP(hipGetLastError());
  return WDL_OK;
}

// ---- the Wasserstein loss (loss.h:92-154) ----------------------------------------
// Its Bregman projections between the barycenter q and the observed o, 50
// iterations from b = 1: c = K b, a = q / c, c = K a, b = o / c (:106-117), run
// on log a, log b with the exact-shift passes (the scalings leave FP32 range at
// small gamma); loss gamma * sum(safe_log(a) q + safe_log(b) o - a K b) (:119-127)
// and gradient gamma * safe_log(a) (:148-150).

__global__ void k_wl_init(SolveArgs a) {
  const long long tot = (long long)a.np * a.N, off = (long long)a.p0 * a.N;
  XLOOP(tot) a.wlb[off + e] = 0.0;
}

// log a = log q - conv (second = false) or log b = log o - conv (second = true)
__global__ void k_wl_update(SolveArgs a, int second) {
  const long long tot = (long long)a.np * a.N, off = (long long)a.p0 * a.N;
  XLOOP(tot) {
    const long long i = off + e;
    const float v = second ? a.obs[i] : a.Q[i];
    const double lv = v > 0.f ? (double)logf(v) : -INFINITY;
    (second ? a.wlb : a.wla)[i] = lv - a.wlc[i];
  }
}

__device__ inline double safe_log_d(double l) { return l > -575.64627324851142 ? l : -575.64627324851142; }

// one workgroup per histogram: loss, gradient and sum of g_L = gradient * q
__global__ __launch_bounds__(kThreads) void k_wl_final(SolveArgs a) {
  const int p = a.p0 + (int)blockIdx.x;
  const long long N = a.N, off = (long long)p * N;
  if (steps_of(a, p) == 0) {
    if (threadIdx.x < kLossBlocks) {
      a.lossPart[(long long)p * kLossBlocks + threadIdx.x] = 0.0;
      a.gsumPart[(long long)p * kLossBlocks + threadIdx.x] = 0.0;
    }
    return;
  }
  double l = 0, gs = 0;
  for (long long x = threadIdx.x; x < N; x += blockDim.x) {
    const double la = safe_log_d(a.wla[off + x]), lb = safe_log_d(a.wlb[off + x]);
    const float q = a.Q[off + x], o = a.obs[off + x];
    const float g = (float)(a.gamma * la);
    a.G0[off + x] = g;
    l += la * q + lb * o - exp(a.wla[off + x] + a.wlc[off + x]);
    gs += (double)(g * q);
  }
  const double tl = block_sum(l), tg = block_sum(gs);
  if (threadIdx.x < kLossBlocks) {
    a.lossPart[(long long)p * kLossBlocks + threadIdx.x] = threadIdx.x == 0 ? a.gamma * tl : 0.0;
    a.gsumPart[(long long)p * kLossBlocks + threadIdx.x] = threadIdx.x == 0 ? tg : 0.0;
  }
}

int wasserstein_loss(const SolveArgs& a, hipStream_t st) {
  const long long n = (long long)a.np * a.N, off = (long long)a.p0 * a.N;
  const int W = a.G.W, H = a.G.H;
  int e;
  hipLaunchKernelGGL(k_wl_init, ew_grid(n), dim3(256), 0, st, a);
  for (int it = 0; it <= kWassersteinIters; ++it) {
    // c = K b; the 51st convolution is the one the loss evaluates (:120)
    if ((e = xlse_conv(a.wlb + off, nullptr, a.wlc + off, nullptr, a.wlt + off, nullptr, W, H,
                       a.np, a.gamma, a.Lp, a.p0, 1, 1, st)))
      return e;
    if (it == kWassersteinIters) break;
    hipLaunchKernelGGL(k_wl_update, ew_grid(n), dim3(256), 0, st, a, 0);
    if ((e = xlse_conv(a.wla + off, nullptr, a.wlc + off, nullptr, a.wlt + off, nullptr, W, H,
                       a.np, a.gamma, a.Lp, a.p0, 1, 1, st)))
      return e;
    hipLaunchKernelGGL(k_wl_update, ew_grid(n), dim3(256), 0, st, a, 1);
  }
  hipLaunchKernelGGL(k_wl_final, dim3((unsigned)a.np), dim3(kThreads), 0, st, a);
  WDL_HIP(hipGetLastError());
  return WDL_OK;
}

#undef XLOOP

}  // namespace wdl
