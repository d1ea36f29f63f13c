// Exact-shift log-domain convolution, and the log-domain solver stages built
// on it (WDL_FLAG_EXACT_LSE).
//
// Each 1-D pass computes, for every output j of a line x,
//     m_j = max(-1e99, max_k(x_k - (j-k)^2/gamma))               (kernels.h:1082-1091)
//     out_j = m_j + log |sum_k s_k exp(x_k - (j-k)^2/gamma - m_j)|
// with the reference's exact per-output maximum, so any spread of log values
// along a line is handled (the shifted MFMA form of wdl_logconv.h needs every
// line within ~80 nats).  Prefix and suffix maxima of the line bound every
// term by distance, which fixes for each output a window outside which all
// terms lie more than kXMargin = 25 nats below m_j (each below e^-25 relative:
// at most len * 1.4e-11 of the sum, under the FP32 rounding of the sum); the
// window gives m_j exactly and holds every term of the sum that counts.  Terms
// are formed in FP64, exponentiated and summed in FP32 (each term <= ~1).  Cost
// per output: the window's terms, ~2 sqrt(25 gamma) + 1 when the log values are
// smooth, up to len when they are not.
//
// The solver stages (exact_forward / exact_backward) run the log-domain
// recursions of compute_gradient_both_log_signArray (inverseWasserstein.h:
// 678-814) as elementwise kernels before the passes and as output stages of
// their second pass, with FP64 scratch images between them.
#include <algorithm>
#include <atomic>
#include <cmath>

#include "wdl_solve_common.h"

namespace wdl {

// Terms outside a window lie this far below m_j: each below e^-25 = 1.4e-11
// relative, at most len * 1.4e-11 <= 1.4e-8 of the sum for lines up to 1024,
// below the FP32 rounding of the sum itself.
constexpr float kXMargin = 25.f;


// Maximum of XF_k - t^2/gamma over k = jc + t, t = -DL..DR.  Without CLAMP
// every k lies in the padded line and |t| <= kXPad, and the penalty comes from
// the table centred on t = 0 (CFs[t], t = -kXPad..kXPad: one address for the
// whole window, stepped by immediate offsets); with it, positions outside the
// line count as -inf and the penalty is formed as the table's entries are,
// (float)(t t) / gamma.
template <bool CLAMP>
__device__ inline float xlse_window_max(const float* XF, const float* CFs, int jc, int len,
                                        int DL, int DR, float igf) {
  auto ld = [&](int t) {
    const int k = jc + t;
    if (CLAMP) return (unsigned)k < (unsigned)len ? XF[k] : -INFINITY;
    return XF[k];
  };
  auto pen = [&](int t) { return CLAMP ? (float)(t * t) * igf : CFs[t]; };
  float m0 = -INFINITY, m1 = -INFINITY;
  int t = -DL;
  for (; t + 3 <= DR; t += 4) {
    float x[4], c[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      x[u] = ld(t + u);
      c[u] = pen(t + u);
    }
    m0 = fmaxf(m0, x[0] - c[0]);
    m1 = fmaxf(m1, x[1] - c[1]);
    m0 = fmaxf(m0, x[2] - c[2]);
    m1 = fmaxf(m1, x[3] - c[3]);
  }
  for (; t <= DR; ++t) m0 = fmaxf(m0, ld(t) - pen(t));
  return fmaxf(m0, m1);
}

// Sum of s_k 2^((X2_k - t^2 log2(e)/gamma) - m2) over the same window (C2s:
// the centred FP64 table; with CLAMP t t (log2(e)/gamma) formed as its entries
// are, ig2 = log2(e)/gamma).
template <bool CLAMP, bool SIGNED>
__device__ inline float xlse_window_sum(const double* X2, const float* SG, const double* C2s,
                                        int jc, int len, int DL, int DR, double m2, double ig2) {
  auto pen = [&](int t) { return CLAMP ? (double)t * t * ig2 : C2s[t]; };
  auto term = [&](double x2, double c, float sg) {
    const float e = __builtin_amdgcn_exp2f((float)((x2 - c) - m2));
    return SIGNED ? sg * e : e;
  };
  auto ldx = [&](int t) {
    const int k = jc + t;
    if (CLAMP) return (unsigned)k < (unsigned)len ? X2[k] : -(double)INFINITY;
    return X2[k];
  };
  auto lds = [&](int t) {
    if (!SIGNED) return 1.f;
    const int k = jc + t;
    if (CLAMP) return (unsigned)k < (unsigned)len ? SG[k] : 0.f;
    return SG[k];
  };
  float s0 = 0.f, s1 = 0.f;
  int t = -DL;
  for (; t + 3 <= DR; t += 4) {
    double x[4], c[4];
    float sg[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      x[u] = ldx(t + u);
      c[u] = pen(t + u);
      sg[u] = lds(t + u);
    }
    s0 += term(x[0], c[0], sg[0]);
    s1 += term(x[1], c[1], sg[1]);
    s0 += term(x[2], c[2], sg[2]);
    s1 += term(x[3], c[3], sg[3]);
  }
  for (; t <= DR; ++t) s0 += term(ldx(t), pen(t), lds(t));
  return s0 + s1;
}

// The wave's window radii {DL, DR}: on each side the largest distance d at
// which some lane's bound B[jc + dir d] - d^2/gamma still exceeds its
// threshold (B: prefix maxima stepping left, dir = -1, suffix maxima stepping
// right, dir = 1), i.e. the widest of the lanes' own radii.  A lane's bound
// falls with d and B holds -inf for kXBPad positions past either end of the
// line, so a lane's passing distances form a prefix ending within the line,
// and so do the distances that pass for any lane: the distances 4, 8, 12, 16
// are tested at once by ballot, further multiples of 4 while all of them
// pass (positions past the line read from its first padding entry), then the
// three distances below the first failing multiple.  Lanes past the line carry
// thr = +inf and pass nowhere.  Penalties are formed as the tables' entries
// are, (float)(d d) / gamma (CFs: the centred table).
struct XRadii {
  int dl, dr;
};
__device__ inline XRadii xlse_wave_radii(const float* PM, const float* SM, const float* CFs,
                                         int jc, int len, float thr, float igf) {
  auto pen = [&](int d) { return (float)(d * d) * igf; };
  // 1 if some lane passes; formed on the 64-bit mask so that it stays scalar
  auto any = [](bool p) {
    const unsigned long long m = __builtin_amdgcn_ballot_w64(p);
    return (int)((m | (0ull - m)) >> 63);
  };
  float r[4], l[4];  // independent reads, issued together
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    r[i] = SM[jc + 4 * i + 4];
    l[i] = PM[jc - 4 * i - 4];
  }
  int cr = 0, cl = 0;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    cr += any(r[i] - pen(4 * i + 4) > thr);
    cl += any(l[i] - pen(4 * i + 4) > thr);
  }
  int br = 4 * cr, bl = 4 * cl;  // pass (or 0, the centre); br + 4, bl + 4 fail
  for (int c = cr; c == 4;) {
    c = 0;
#pragma unroll
    for (int i = 1; i <= 4; ++i) c += any(SM[min(jc + br + 4 * i, len)] - pen(br + 4 * i) > thr);
    br += 4 * c;
  }
  for (int c = cl; c == 4;) {
    c = 0;
#pragma unroll
    for (int i = 1; i <= 4; ++i) c += any(PM[max(jc - bl - 4 * i, -1)] - pen(bl + 4 * i) > thr);
    bl += 4 * c;
  }
  // the three distances below the first failing multiple of 4: within the
  // padding and the penalty table when it is at most 16
  float r3[3], l3[3], pr[3], pl[3];
  if (br <= 12 && bl <= 12) {
#pragma unroll
    for (int u = 0; u < 3; ++u) {
      r3[u] = SM[jc + br + 1 + u];
      l3[u] = PM[jc - bl - 1 - u];
      pr[u] = CFs[br + 1 + u];
      pl[u] = CFs[bl + 1 + u];
    }
  } else {
#pragma unroll
    for (int u = 0; u < 3; ++u) {
      r3[u] = SM[min(jc + br + 1 + u, len)];
      l3[u] = PM[max(jc - bl - 1 - u, -1)];
      pr[u] = pen(br + 1 + u);
      pl[u] = pen(bl + 1 + u);
    }
  }
  XRadii w{bl, br};
#pragma unroll
  for (int u = 0; u < 3; ++u) {
    w.dr += any(r3[u] - pr[u] > thr);
    w.dl += any(l3[u] - pl[u] > thr);
  }
  return w;
}

constexpr int kXMaxThreads = 1024;

// One pass over `nlines` lines of length `len` (line-major input), writing the
// transposed layout (line index fastest) of each image of `rows` lines.  A line
// is skipped (output untouched) when active_thr > 0 and Lp[p0 + image / ipp] <
// active_thr.
//
// LW lines per workgroup, xlse_waves_per_line(len) waves each; one output
// j per lane, 64 consecutive outputs at a time.  Every term of output j is
// x_k - (j-k)^2/gamma <= PM[k] - (j-k)^2/gamma for k <= j (PM: prefix maxima)
// and <= SM[k] - (k-j)^2/gamma for k >= j (SM: suffix maxima), and both bounds
// fall as k moves away from j.  With m0 the maximum of the five terms
// |j-k| <= 2, the lane's window ends on each side before the first distance d
// where the bound at j +- d is at most m0 - kXMargin: every term beyond lies
// more than kXMargin nats below m0 <= m_j.  The 64 lanes share one window [j-DL, j+DR]
// (the widest of theirs), so the loops over it run in step with the penalty
// t^2/gamma read from one LDS table.  The window gives m_j exactly (terms
// outside it are below m0); the sum adds every term within it, so the terms
// left out are each below e^-kXMargin relative to the largest.  m_j is found in FP32
// (any shift near the maximum serves: the output adds back the shift the terms
// were formed with); exponents are formed in FP64 in log2 units and
// exponentiated on the hardware exp2.
//
// blk / nblk: this workgroup's index and the number of workgroups sharing the
// lines (the grid, or a member's share of a group launch).
template <bool SIGNED, class Epi, int LW, int KP = 4>
__device__ __forceinline__ void xlse_pass_body(const double* __restrict__ in,
                                               const unsigned char* __restrict__ sin, Epi epi,
                                               int rows, int len, long long nlines,
                                               double inv_gamma, const int* __restrict__ Lp,
                                               int p0, int ipp, int active_thr, int blk,
                                               int nblk) {
  extern __shared__ __align__(16) double xs[];
  const int lp = len + 2 * kXPad;
  const int wpl = blockDim.x / (LW * 64);  // waves per line
  const int wave = wave_id(), lane = threadIdx.x & 63;
  const int w = wave % LW, sub = wave / LW;  // line slot, wave within the line
  const float igf = (float)inv_gamma;
  const double kLog2e = 1.4426950408889634;
  // layout: X2 [LW][lp], O0 [LW][len], C2s [kXSym] (doubles);
  // XF, SG [LW][lp], PM, SM [LW][lb], CFs [kXSym] (floats);
  // SO0 [LW][len]
  double* X2 = xs + w * lp + kXPad;
  double* O0 = xs + LW * lp;
  double* O = O0 + w * len;
  double* C2s = O0 + LW * len + kXPad;  // C2s[t], t = -kXPad..kXPad
  float* F0 = reinterpret_cast<float*>(C2s + kXPad + 1);
  float* XF = F0 + w * lp + kXPad;
  float* SG = F0 + LW * lp + w * lp + kXPad;
  const int lb = len + 2 * kXBPad;
  float* PM = F0 + 2 * LW * lp + w * lb + kXBPad;
  float* SM = F0 + 2 * LW * lp + LW * lb + w * lb + kXBPad;
  float* CFs = F0 + 2 * LW * lp + 2 * LW * lb + kXPad;
  unsigned char* SO0 = reinterpret_cast<unsigned char*>(CFs + kXPad + 1);
  unsigned char* SO = SO0 + w * len;
  const double ig2 = inv_gamma * kLog2e;
  for (int t = threadIdx.x; t < kXSym; t += blockDim.x) {
    const int d = t - kXPad;
    C2s[d] = (double)d * d * ig2;
    CFs[d] = (float)(d * d) * igf;
  }
  // padding: the window loops read up to kXPad positions past either end,
  // the window search up to kXBPad
  if (sub == 0 && lane < kXPad) {
    X2[-1 - lane] = X2[len + lane] = -(double)INFINITY;
    XF[-1 - lane] = XF[len + lane] = -INFINITY;
    SG[-1 - lane] = SG[len + lane] = 0.f;
  }
  if (sub == 0 && lane < kXBPad) PM[-1 - lane] = PM[len + lane] = SM[-1 - lane] = SM[len + lane] = -INFINITY;
  // Per line slot: whether its line runs (1) or is skipped, for the store.
  unsigned char* ACT = SO0 + LW * len;
  // The next line group's values (at most KP per lane: 4 for lines up to 1024
  // over 4 waves, 2 where a line has at most 128 positions per wave), sign
  // bytes and iteration counts, loaded one group ahead so
  // the loads' latency passes under the current group's work.  Lane i loads
  // the count of line slot i % LW and the wave reads its slot's lane when the
  // group starts: a load at the wave's own (uniform) address would be a
  // scalar load, waited for at the next LDS access (both count on the same
  // counter), i.e. at once.
  constexpr int kXPer = KP;
  double pv[kXPer];
  unsigned char ps[kXPer];
  bool pinb = false;
  int plp = 0;
  // image of line ln: in 32-bit arithmetic when every line index fits
  const bool lines32 = nlines <= 0xffffffffLL;
  auto image_of = [&](long long ln) -> long long {
    return lines32 ? (long long)((unsigned)ln / (unsigned)rows) : ln / rows;
  };
  auto fetch = [&](long long l0n) {
    const long long ln = l0n + w;
    const bool inb = ln < nlines;
    pinb = inb;
    if (active_thr > 0) {
      const long long li = l0n + lane % LW;
      if (li < nlines) plp = Lp[p0 + (int)image_of(li) / ipp];
    }
#pragma unroll
    for (int u = 0; u < kXPer; ++u) {
      const int k = sub * 64 + lane + u * wpl * 64;
      if (inb && k < len) {
        pv[u] = in[ln * len + k];
        if (SIGNED) ps[u] = sin[ln * len + k];
      }
    }
  };
  const long long stride = (long long)nblk * LW;
  long long l0 = (long long)blk * LW;
  if (l0 < nlines) fetch(l0);
  for (; l0 < nlines; l0 += stride) {
    const int lpw = __builtin_amdgcn_readlane(plp, w);
    const bool active = pinb && (active_thr <= 0 || lpw >= active_thr);
    if (active) {  // stage the line: 64 values per wave and step
#pragma unroll
      for (int u = 0; u < kXPer; ++u) {
        const int k = sub * 64 + lane + u * wpl * 64;
        if (k < len) {
          X2[k] = pv[u] * kLog2e;
          XF[k] = (float)pv[u];
          if (SIGNED) SG[k] = ps[u] ? 1.f : -1.f;  // getSign (signArray.h:19-22)
        }
      }
    }
    if (sub == 0 && lane == 0) ACT[w] = active ? 1 : 0;
    __syncthreads();
    if (l0 + stride < nlines) fetch(l0 + stride);
    // prefix maxima (wave 0 of the line) and suffix maxima (its last wave):
    // inclusive max-scans across the wave (wave_scan), 64 positions per
    // step, with a running carry from lane 63; a line of one wave runs both
    // in one loop, so that the two chains of dependent steps interleave
    if (active && wpl == 1) {
      float cp = -INFINITY, cs = -INFINITY;
      for (int base = 0; base < len; base += 64) {
        const int k = base + lane, q = len - 1 - base - lane;
        const float v = fmaxf(wave_scan(k < len ? XF[k] : -INFINITY, MaxOp()), cp);
        const float u = fmaxf(wave_scan(q >= 0 ? XF[q] : -INFINITY, MaxOp()), cs);
        if (k < len) PM[k] = v;
        if (q >= 0) SM[q] = u;
        cp = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
        cs = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(u), 63));
      }
    } else if (active && sub == 0) {
      float carry = -INFINITY;
      for (int base = 0; base < len; base += 64) {
        const int k = base + lane;
        float v = fmaxf(wave_scan(k < len ? XF[k] : -INFINITY, MaxOp()), carry);
        if (k < len) PM[k] = v;
        carry = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
      }
    }
    if (active && wpl > 1 && sub == wpl - 1) {
      float carry = -INFINITY;
      for (int top = len - 1; top >= 0; top -= 64) {
        const int k = top - lane;
        float v = fmaxf(wave_scan(k >= 0 ? XF[k] : -INFINITY, MaxOp()), carry);
        if (k >= 0) SM[k] = v;
        carry = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
      }
    }
    __syncthreads();
    if (active)
      for (int j0 = sub * 64; j0 < len; j0 += wpl * 64) {
        const int j = j0 + lane;
        const bool valid = j < len;
        const int jc = valid ? j : len - 1;
        float m0 = XF[jc];
#pragma unroll
        for (int t = 1; t <= 2; ++t) m0 = fmaxf(m0, fmaxf(XF[jc - t], XF[jc + t]) - CFs[t]);
        const XRadii wr =
            xlse_wave_radii(PM, SM, CFs, jc, len, valid ? m0 - kXMargin : INFINITY, igf);
        const int DL = wr.dl, DR = wr.dr;
        const bool clamp = DL > kXPad || DR > kXPad;
        const float m = clamp ? xlse_window_max<true>(XF, CFs, jc, len, DL, DR, igf)
                              : xlse_window_max<false>(XF, CFs, jc, len, DL, DR, igf);
        const double m64 = fmax((double)m, -1e99);
        const double m2 = m64 * kLog2e;
        const float sum =
            clamp ? xlse_window_sum<true, SIGNED>(X2, SG, C2s, jc, len, DL, DR, m2, ig2)
                  : xlse_window_sum<false, SIGNED>(X2, SG, C2s, jc, len, DL, DR, m2, ig2);
        if (valid) {
          O[j] = m64 + hlog(fabsf(sum));
          if (SIGNED) SO[j] = sum >= 0.f ? 1 : 0;
        }
      }
    __syncthreads();
    // transposed store: the lines of one workgroup are consecutive, so each
    // output position j gets up to LW adjacent values.  Thread i serves
    // line slot i % LW throughout (blockDim is a multiple of LW), so the
    // line's image, offset and output stage (epi.line: its pointers and
    // per-image values) are worked out once per line group.
    {
      const int ws = threadIdx.x % LW;
      const long long ln = l0 + ws;
      if (ACT[ws] != 0) {  // implies ln < nlines
        const long long img = image_of(ln);
        const int r = (int)(ln - img * rows);
        const auto out = epi.line(img, (long long)rows * len);
        for (int j = threadIdx.x / LW; j < len; j += blockDim.x / LW)
          out(r + j * rows, O0[ws * len + j], SIGNED ? SO0[ws * len + j] : 1);
      }
    }
    __syncthreads();
  }
}

// Waves per SIMD.  KP: values per lane prefetched (2 when each wave of a line
// holds at most 128 positions).  With 2, 7 for the unsigned passes (71 VGPRs)
// and 6 for the signed ones (79 VGPRs; the backward output stages spill 12
// bytes): each ~1% more Sinkhorn-iters/s on the 512-request 128x128
// projection than 6 and 5.  With 4, 5: the signed passes would take ~100 VGPRs
// and hold 4; limited to 96 they spill a few prefetch addresses, once per line
// group (59.1k against 57.8k on the same projection).
template <bool SIGNED, class Epi, int LW, int KP>
__global__ __launch_bounds__(kXMaxThreads) __attribute__((amdgpu_waves_per_eu(KP > 2 ? 5 : (SIGNED ? 6 : 7)))) void k_xlse_pass(
    const double* __restrict__ in, const unsigned char* __restrict__ sin, Epi epi, int rows,
    int len, long long nlines, double inv_gamma, const int* __restrict__ Lp, int p0, int ipp,
    int active_thr) {
  xlse_pass_body<SIGNED, Epi, LW, KP>(in, sin, epi, rows, len, nlines, inv_gamma, Lp, p0, ipp,
                                      active_thr, blockIdx.x, gridDim.x);
}

// Output stage of k_xlse_pass: line(img, N) gives, for image img (of N
// values) of the pass, the stage of that image's outputs, called once per
// output (pixel x, value, sign byte).  The plain stage stores them.
struct XStore {
  double* out;
  unsigned char* sout;  // nullptr without signs
  struct Line {
    double* o;
    unsigned char* so;
    __device__ void operator()(int x, double v, unsigned char sg) const {
      o[x] = v;
      if (so) so[x] = sg;
    }
  };
  __device__ Line line(long long img, long long N) const {
    return Line{out + img * N, sout ? sout + img * N : nullptr};
  }
};

// Compute units of the device, for the pass's grid (set with the LDS attributes).
static std::atomic<int> g_xlse_cus{0};

// Solvers may be created from several host threads (the server's included);
// a repeated attribute call is harmless, so a relaxed atomic suffices.
template <bool SIGNED, class Epi, int LW, int KP>
static int xlse_set_attr_lw(int lds) {
  static std::atomic<int> done{0};
  if (lds <= done.load(std::memory_order_relaxed)) return WDL_OK;
  WDL_HIP(hipFuncSetAttribute((const void*)k_xlse_pass<SIGNED, Epi, LW, KP>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, lds));
  int cur = done.load(std::memory_order_relaxed);
  while (cur < lds && !done.compare_exchange_weak(cur, lds, std::memory_order_relaxed)) {
  }
  return WDL_OK;
}


// maxlen: the longest line the passes will run (the LDS attribute to allow).
template <bool SIGNED, class Epi>
static int xlse_set_attr(int maxlen) {
  int e;
  const int l4 = xlse_lds_bytes(maxlen, 4), l8 = xlse_lds_bytes(maxlen, 8);
  if ((e = xlse_set_attr_lw<SIGNED, Epi, 4, 4>(l4)) || (e = xlse_set_attr_lw<SIGNED, Epi, 4, 2>(l4)))
    return e;
  if (l8 > kXLdsMax) return WDL_OK;
  if ((e = xlse_set_attr_lw<SIGNED, Epi, 8, 4>(l8))) return e;
  return xlse_set_attr_lw<SIGNED, Epi, 8, 2>(l8);
}

// One pass over nimg images of rows lines of length len (see k_xlse_pass).
template <bool SIGNED, class Epi>
static void xlse_launch(const double* in, const unsigned char* sin, Epi epi, int rows, int len,
                        long long nimg, double ig, const int* Lp, int p0, int ipp, int thr,
                        int wpl, int lw, hipStream_t st) {
  const long long nl = nimg * rows;
  const int w = xlse_waves_per_line(len, wpl);
  lw = xlse_lines(len, w, lw);
  const int lds = xlse_lds_bytes(len, lw);
  // twice the workgroups the CUs hold at once, each looping over line groups
  const int perCU = std::max(1, std::min(160 * 1024 / lds, 8));
  const int cus = g_xlse_cus.load(std::memory_order_relaxed);
  const long long cap = 2LL * (cus > 0 ? cus : 256) * perCU;
  const int grid = (int)std::min<long long>((nl + lw - 1) / lw, cap);
  const int block = lw * 64 * w;
  const bool kp2 = len <= 2 * 64 * w;  // at most 2 values per lane to prefetch
#define WDL_XLAUNCH(LW_, KP_)                                                                   \
  hipLaunchKernelGGL((k_xlse_pass<SIGNED, Epi, LW_, KP_>), dim3(grid), dim3(block), lds, st, in, \
                     sin, epi, rows, len, nl, ig, Lp, p0, ipp, thr)
  if (lw == 8) {
    if (kp2) WDL_XLAUNCH(8, 2); else WDL_XLAUNCH(8, 4);
  } else {
    if (kp2) WDL_XLAUNCH(4, 2); else WDL_XLAUNCH(4, 4);
  }
#undef WDL_XLAUNCH
}

// Two passes (along x, then along y) over nimg W x H images; tmp holds the
// transposed intermediate.  Signs: none (sin == nullptr) or one byte per value.
// The second pass hands its outputs to epi (XStore: into out / sout).
template <class Epi>
static int xlse_conv_e(const double* in, const unsigned char* sin, Epi epi, double* tmp,
                       unsigned char* stmp, int W, int H, long long nimg, double gamma,
                       const int* Lp, int p0, int ipp, int thr, hipStream_t st, int wpl = 0,
                       int lw = kXWaves) {
  const double ig = 1.0 / gamma;
  const int thr_ = Lp ? thr : 0;
  if (sin) {
    xlse_launch<true>(in, sin, XStore{tmp, stmp}, H, W, nimg, ig, Lp, p0, ipp, thr_, wpl, lw, st);
    xlse_launch<true>(tmp, stmp, epi, W, H, nimg, ig, Lp, p0, ipp, thr_, wpl, lw, st);  // columns
  } else {
    xlse_launch<false>(in, nullptr, XStore{tmp, nullptr}, H, W, nimg, ig, Lp, p0, ipp, thr_, wpl,
                       lw, st);
    xlse_launch<false>(tmp, nullptr, epi, W, H, nimg, ig, Lp, p0, ipp, thr_, wpl, lw, st);
  }
  WDL_HIP(hipGetLastError());
  return WDL_OK;
}

static int xlse_conv(const double* in, const unsigned char* sin, double* out, unsigned char* sout,
                     double* tmp, unsigned char* stmp, int W, int H, long long nimg, double gamma,
                     const int* Lp, int p0, int ipp, int thr, hipStream_t st, int wpl = 0,
                     int lw = kXWaves) {
  return xlse_conv_e(in, sin, XStore{out, sin ? sout : nullptr}, tmp, stmp, W, H, nimg, gamma, Lp,
                     p0, ipp, thr, st, wpl, lw);
}

// Solver output stages fused into the second pass (the image layout of the
// scratch arrays: image img = (histogram - p0, chain, atom), pixel x).

// forward, first convolution: c^{s-1} into the history and the second
// convolution's input log p - c^{s-1} (:704-707)
struct XEpiFwd1 {
  double* ch;  // Ch[s-1] at the first image
  double* xa;  // the solver's xa at the first image
  const float* logA;
  int K;
  long long N;
  struct Line {
    double* ch;
    double* xa;
    const float* la;
    __device__ void operator()(int x, double c, unsigned char) const {
      ch[x] = c;
      xa[x] = (double)la[x] - c;
    }
  };
  __device__ Line line(long long img, long long) const {
    return Line{ch + img * N, xa + img * N, logA + (long long)((int)img % K) * N};
  }
};

// The fields of SolveArgs the backward output stages read: the stages carry
// this instead of the whole SolveArgs, which a group launch would otherwise
// copy out of its member table into scratch memory.
struct XBwdArgs {
  long long N, stage;
  int K, p0, L, ring;
  const int* Lp;
  const float *Ch, *Fh, *Bh, *logA;
  float *GD, *V0, *V1, *R0, *R1;
  double* xa;
  unsigned char* sga;
};
__host__ __device__ inline XBwdArgs xbwd_args(const SolveArgs& a) {
  return XBwdArgs{a.N,  a.stage, a.K,  a.p0, a.L,  a.ring, a.Lp,  a.Ch,  a.Fh,
                  a.Bh, a.logA,  a.GD, a.V0, a.V1, a.R0,   a.R1,  a.xa,  a.sga};
}

// backward step s, after the first convolution: resultDic (chain 0, :750-752)
// and, for s >= 2, the second convolution's input log p + c - 2 log c^{s-1},
// sign of c (:754-756, :796)
struct XEpiBwd1 {
  XBwdArgs a;
  int t, nch;
  long long off;  // scratch offset of the first image
  struct Line {
    const double* cs;  // log c^{s-1} of the image
    float* gd;         // resultDic of the image (chain 0), else nullptr
    double* xa;
    unsigned char* sga;
    const float* la;   // log p of the atom
    bool first;        // the histogram's last step: resultDic starts at 0
    bool more;         // s >= 2: the second convolution follows
    __device__ void operator()(int x, double c, unsigned char sgb) const {
      const double csx = cs[x];
      if (gd) {
        const float prev = first ? 0.f : gd[x];
        gd[x] = prev + (sgb ? 1.f : -1.f) * expf((float)(c - csx));
      }
      if (more) {
        xa[x] = (double)la[x] + c - 2.0 * csx;
        sga[x] = sgb;
      }
    }
  };
  __device__ Line line(long long img, long long) const {
    const long long N = a.N;
    const int im = (int)img, i = im % a.K, pc = im / a.K;  // images: fewer than 2^31
    const int chain = nch == 1 ? 1 : pc % nch;
    const int p = a.p0 + pc / nch;
    const int s = bwd_step(a, p, t);
    const long long pimg = ((long long)p * a.K + i) * N;
    Line l;
    l.cs = (const double*)a.Ch + hslot(a, s - 1) * a.stage + pimg;
    l.gd = chain == 0 ? a.GD + pimg : nullptr;
    l.first = s == steps_of(a, p);
    l.more = s > 1;
    l.xa = a.xa + off + img * N;
    l.sga = a.sga + off + img * N;
    l.la = a.logA + (long long)i * N;
    return l;
  }
};

// backward step s >= 2, after the second convolution: -v = sign exp(res - log
// phi^{s-1}) (:757-762), -r = sign exp(res + log b^{s-1}) (:797-805)
struct XEpiBwd2 {
  XBwdArgs a;
  int t, nch;
  struct Line {
    const double* src;  // log phi^{s-1} (chain 0) or log b^{s-1} (chain 1) of the image
    float* dst;         // -v or -r of the image
    double sgn;         // -1 (chain 0: res - log phi) or +1 (chain 1: res + log b)
    __device__ void operator()(int x, double v, unsigned char sgb) const {
      dst[x] = (sgb ? 1.f : -1.f) * expf((float)(v + sgn * src[x]));
    }
  };
  __device__ Line line(long long img, long long) const {
    const long long N = a.N;
    const int im = (int)img, i = im % a.K, pc = im / a.K;  // images: fewer than 2^31
    const int chain = nch == 1 ? 1 : pc % nch;
    const int p = a.p0 + pc / nch;
    const int s = bwd_step(a, p, t);
    const long long pimg = ((long long)p * a.K + i) * N;
    const long long h = hslot(a, s - 2) * a.stage + pimg;
    if (chain == 0) return Line{(const double*)a.Fh + h, ((s & 1) ? a.V1 : a.V0) + pimg, -1.0};
    return Line{(const double*)a.Bh + h, ((s & 1) ? a.R1 : a.R0) + pimg, 1.0};
  }
};

int xlse_set_lds_attr(int maxlen) {
  int e;
  if (g_xlse_cus.load(std::memory_order_relaxed) == 0) {
    int dev = 0, cus = 0;
    WDL_HIP(hipGetDevice(&dev));
    WDL_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    g_xlse_cus.store(std::max(cus, 1), std::memory_order_relaxed);
  }
  if ((e = xlse_set_attr<true, XStore>(maxlen)) || (e = xlse_set_attr<false, XStore>(maxlen)) ||
      (e = xlse_set_attr<false, XEpiFwd1>(maxlen)) || (e = xlse_set_attr<true, XEpiBwd1>(maxlen)) ||
      (e = xlse_set_attr<true, XEpiBwd2>(maxlen)))
    return e;
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

// Grid-stride loop over a kernel's elements; blk_ / nblk_ are the workgroup's
// index and the workgroups sharing the elements (the grid, or a member's share
// of a group launch).
#define XLOOP(total)                                                                     \
  for (long long e = (long long)blk_ * blockDim.x + threadIdx.x; e < (total);            \
       e += (long long)nblk_ * blockDim.x)
#define XGRID const int blk_ = blockIdx.x, nblk_ = gridDim.x

// forward, input of the first convolution: log b^{s-1} (:711-724; b^0: 0 or the store)
__device__ inline void xf1_body(const SolveArgs& a, int s, int blk_, int nblk_) {
  // one thread per (histogram, pixel): the geometric mean of the K atoms'
  // phi^{s-1} is formed once and gives all K images' b
  const long long N = a.N, tot = (long long)a.np * N;
  XLOOP(tot) {
    const int p = a.p0 + (int)(e / N);
    const long long x = e % N;
    if (!fwd_runs(a, p, s)) continue;
    const long long hb = (long long)p * a.K * N;
    if (s == 1) {
      for (int i = 0; i < a.K; ++i)
        a.xa[a.xfwd + hb + i * N + x] = a.lbInit ? a.lbInit[hb + i * N + x] : 0.0;
    } else {
      const double* Fp = (const double*)a.Fh + hslot(a, s - 2) * a.stage + hb + x;
      double* Bp = (double*)a.Bh + hslot(a, s - 2) * a.stage + hb + x;
      double lq = 0;
      const float* lam = a.lam + (long long)p * a.K;
      for_atoms_l<4>(
          a.K, [&](int j) { return Fp[j * N]; },
          [&](int j, double f) { lq += (double)lam[j] * f; });
      for (int i = 0; i < a.K; ++i) {
        const double b = lq - Fp[i * N];
        Bp[i * N] = b;
        a.xa[a.xfwd + hb + i * N + x] = b;  // image (p, i) of the forward's layout
      }
    }
  }
}

__global__ void k_xf1(SolveArgs a, int s) { xf1_body(a, s, blockIdx.x, gridDim.x); }

// backward step, first convolution's input: chain 0 (lam_i n + v_i) b^s, chain 1
// (lam_i g + r_i) / phi^s, as log magnitude and sign (:740-749, :778-793)
__device__ inline void xb1_body(const SolveArgs& a, int t, int nch, int blk_, int nblk_) {
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
    const long long h = hslot(a, s - 1) * a.stage + r.img + r.x;
    const double lh = chain == 0 ? ((const double*)a.Bh)[h] : -((const double*)a.Fh)[h];
    const long long o = xoff(a, nch, r);
    a.xa[o] = hlog(fabsf(xv)) + lh;
    a.sga[o] = xv >= 0.f ? 1 : 0;
  }
}

__global__ void k_xb1(SolveArgs a, int t, int nch) { xb1_body(a, t, nch, blockIdx.x, gridDim.x); }

// Weight-gradient dot of backward step s, <log phi_i^s, g_s> (:786), centred as
// in the other solvers (exact_gsum), in two kernels.  k_xdot_chunk: one
// workgroup per (histogram, atom, chunk of kXDotChunk pixels) writes the
// chunk's sums of log phi * g, log phi and g to xdp; with FUSED (the
// weights-only backward, one chain) it also forms the first convolution's
// input as k_xb1 does for the weight chain (:800-808), from the same loads of
// the K atoms' r and of log phi_i^s.  k_xdot_fin adds the chunks in order, so
// both forms give the same dots to the bit.
template <bool FUSED>
__device__ inline void xdot_chunk_body(const SolveArgs& a, int t, int blk) {
  const long long N = a.N, nc = xdot_chunks(N);
  const long long ic = blk / nc;
  const int c = (int)(blk - ic * nc);
  const int p = a.p0 + (int)(ic / a.K), i = (int)(ic % a.K);
  const int s = bwd_step(a, p, t);
  if (s < 1) return;
  const bool last = s == steps_of(a, p);
  const long long hb = (long long)p * a.K * N, img = hb + (long long)i * N;
  const float lpi = a.lam[(long long)p * a.K + i];
  const float* Rp = (s & 1) ? a.R0 : a.R1;
  const double* F = (const double*)a.Fh + hslot(a, s - 1) * a.stage + img;
  const long long xe = (long long)(c + 1) * kXDotChunk, x1 = xe < N ? xe : N;
  double sfg = 0, sf = 0, sg = 0;
  for (long long x = (long long)c * kXDotChunk + threadIdx.x; x < x1; x += blockDim.x) {
    float g = 0.f, own = 0.f;
    if (last) {
      g = a.G0[(long long)p * N + x] * a.Q[(long long)p * N + x];
    } else {
      for_atoms_l(
          a.K, [&](int j) { return Rp[hb + j * N + x]; },
          [&](int j, float v) {
            g -= v;
            if (j == i) own = v;
          });
    }
    const double f = F[x];
    if (FUSED) {
      const float xv = lpi * g + own;
      a.xa[img + x] = hlog(fabsf(xv)) - f;  // image (p, i) of the one-chain layout
      a.sga[img + x] = xv >= 0.f ? 1 : 0;
    }
    sfg += f * (double)g;
    sf += f;
    sg += (double)g;
  }
  const double tfg = block_sum(sfg), tf = block_sum(sf), tg = block_sum(sg);
  if (threadIdx.x == 0) {
    double* o = a.xdp + (((long long)p * a.K + i) * nc + c) * 3;
    o[0] = tfg;
    o[1] = tf;
    o[2] = tg;
  }
}

template <bool FUSED>
__global__ __launch_bounds__(kThreads) void k_xdot_chunk(SolveArgs a, int t) {
  xdot_chunk_body<FUSED>(a, t, blockIdx.x);
}

__device__ inline void xdot_fin_body(const SolveArgs& a, int t, int blk_, int nblk_) {
  const long long N = a.N, nc = xdot_chunks(N);
  const long long tot = (long long)a.np * a.K;
  XLOOP(tot) {
    const int p = a.p0 + (int)(e / a.K), i = (int)(e % a.K);
    const int s = bwd_step(a, p, t);
    if (s < 1) continue;
    const double* q = a.xdp + ((long long)p * a.K + i) * nc * 3;
    double tfg = 0, tf = 0, tg = 0;
    for (long long c = 0; c < nc; ++c) {
      tfg += q[3 * c];
      tf += q[3 * c + 1];
      tg += q[3 * c + 2];
    }
    double* gl = a.GL + (long long)p * a.K + i;
    *gl = s == steps_of(a, p) ? tfg : *gl + (tfg - tf / (double)N * (tg - exact_gsum(a, p, s)));
  }
}

__global__ void k_xdot_fin(SolveArgs a, int t) { xdot_fin_body(a, t, blockIdx.x, gridDim.x); }

static dim3 ew_grid(long long n) {
  return dim3((unsigned)std::min<long long>((n + 255) / 256, 8192));
}

int exact_forward(const SolveArgs& a, hipStream_t st, int s) {
  const long long nimg = (long long)a.np * a.K;
  const long long off = (long long)a.p0 * a.K * a.N;
  const int W = a.G.W, H = a.G.H;
  const double g = a.gamma;
  int e;
  hipLaunchKernelGGL(k_xf1, ew_grid((long long)a.np * a.N), dim3(256), 0, st, a, s);
  // c^{s-1}, and log p - c^{s-1} as the second convolution's input
  double* xa = a.xa + a.xfwd + off;
  double* xt = a.xt + a.xfwd + off;
  const XEpiFwd1 epi{(double*)a.Ch + hslot(a, s - 1) * a.stage + off, xa, a.logA, a.K,
                     a.N};
  const int thr = std::max(s, a.rmin);  // lines of histograms that run iteration s
  if ((e = xlse_conv_e(xa, nullptr, epi, xt, nullptr, W, H, nimg, g, a.Lp, a.p0,
                       a.K, thr, st, a.xwpl, a.xlw)))
    return e;
  if ((e = xlse_conv(xa, nullptr, (double*)a.Fh + hslot(a, s - 1) * a.stage + off,
                     nullptr, xt, nullptr, W, H, nimg, g, a.Lp, a.p0, a.K, thr, st, a.xwpl, a.xlw)))
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
  const dim3 gdot((unsigned)((long long)a.np * a.K * xdot_chunks(a.N)));
  if (a.weightsOnly) {
    hipLaunchKernelGGL(k_xdot_chunk<true>, gdot, dim3(kThreads), 0, st, a, t);
  } else {
    hipLaunchKernelGGL(k_xdot_chunk<false>, gdot, dim3(kThreads), 0, st, a, t);
    hipLaunchKernelGGL(k_xb1, ew_grid(n), dim3(256), 0, st, a, t, nch);
  }
  hipLaunchKernelGGL(k_xdot_fin, ew_grid((long long)a.np * a.K), dim3(256), 0, st, a, t);
  // histogram p runs step t when L_p >= t (the second convolution for t >= 2)
  if ((e = xlse_conv_e(a.xa + off, a.sga + off, XEpiBwd1{xbwd_args(a), t, nch, off}, a.xt + off, a.sgt + off,
                       W, H, nimg, g, a.Lp, a.p0, nch * a.K, t, st, a.xwpl, a.xlw)))
    return e;
  if (t >= 2 && (e = xlse_conv_e(a.xa + off, a.sga + off, XEpiBwd2{xbwd_args(a), t, nch}, a.xt + off,
                                 a.sgt + off, W, H, nimg, g, a.Lp, a.p0, nch * a.K, t, st,
                                 a.xwpl, a.xlw)))
    return e;
  WDL_HIP(hipGetLastError());
  return WDL_OK;
}

// ---- group launches -------------------------------------------------------------
// Each kernel below finds its workgroup's member (group_member), copies the
// member's arguments and runs the single-solver kernel's body on the member's
// share of the grid, with the arguments exact_forward / exact_backward pass,
// formed here from the member's SolveArgs (p0 = 0: a member runs all its
// histograms).  A member whose histograms have fewer iterations than the
// current step (lmax: the largest of its per-histogram counts) returns at
// once, as all its lines and elements would be skipped.

__global__ void k_xf1_group(GroupArgs g, int s) {
  int lb, nb;
  const GroupMember& gm = g.m[group_member(g, GK_XF1, blockIdx.x, lb, nb)];
  if (*gm.lmax < s) return;
  const SolveArgs& a = gm.a;
  xf1_body(a, s, lb, nb);
}

// Forward passes of iteration s: STAGE 0 / 1 the first convolution along x /
// y (output stage XEpiFwd1), 2 / 3 the second (into the history Fh).
template <int STAGE>
__global__ __launch_bounds__(kXMaxThreads) __attribute__((amdgpu_waves_per_eu(5))) void k_xlse_group_fwd(GroupArgs g, int s) {
  int lb, nb;
  const GroupMember& gm =
      g.m[group_member(g, (STAGE & 1) ? GK_FB : GK_FA, blockIdx.x, lb, nb)];
  const SolveArgs& a = gm.a;
  const double ig = gm.ig;
  const int thr = max(s, a.rmin);
  if (*gm.lmax < thr) return;
  const int W = a.G.W, H = a.G.H;
  const long long nimg = (long long)a.np * a.K, off = (long long)a.p0 * a.K * a.N;
  double* xa = a.xa + a.xfwd + off;
  double* xt = a.xt + a.xfwd + off;
  if constexpr (STAGE == 0 || STAGE == 2) {
    xlse_pass_body<false, XStore, kXWaves>(xa, nullptr, XStore{xt, nullptr}, H, W, nimg * H, ig,
                                           a.Lp, a.p0, a.K, thr, lb, nb);
  } else if constexpr (STAGE == 1) {
    const XEpiFwd1 epi{(double*)a.Ch + hslot(a, s - 1) * a.stage + off, xa, a.logA, a.K, a.N};
    xlse_pass_body<false, XEpiFwd1, kXWaves>(xt, nullptr, epi, W, H, nimg * W, ig, a.Lp, a.p0,
                                             a.K, thr, lb, nb);
  } else {
    const XStore epi{(double*)a.Fh + hslot(a, s - 1) * a.stage + off, nullptr};
    xlse_pass_body<false, XStore, kXWaves>(xt, nullptr, epi, W, H, nimg * W, ig, a.Lp, a.p0, a.K,
                                           thr, lb, nb);
  }
}

// Backward passes of step t: STAGE 0 / 1 the first convolution along x / y
// (output stage XEpiBwd1), 2 / 3 the second (XEpiBwd2; launched for t >= 2).
template <int STAGE>
__global__ __launch_bounds__(kXMaxThreads) __attribute__((amdgpu_waves_per_eu(5))) void k_xlse_group_bwd(GroupArgs g, int t) {
  int lb, nb;
  const GroupMember& gm =
      g.m[group_member(g, (STAGE & 1) ? GK_BB : GK_BA, blockIdx.x, lb, nb)];
  const SolveArgs& a = gm.a;
  const double ig = gm.ig;
  if (*gm.lmax < t) return;
  const int W = a.G.W, H = a.G.H, nch = a.weightsOnly ? 1 : 2;
  const long long nimg = (long long)a.np * nch * a.K, off = (long long)a.p0 * nch * a.K * a.N;
  if constexpr (STAGE == 0 || STAGE == 2) {
    xlse_pass_body<true, XStore, kXWaves>(a.xa + off, a.sga + off, XStore{a.xt + off, a.sgt + off},
                                          H, W, nimg * H, ig, a.Lp, a.p0, nch * a.K, t, lb, nb);
  } else if constexpr (STAGE == 1) {
    xlse_pass_body<true, XEpiBwd1, kXWaves>(a.xt + off, a.sgt + off, XEpiBwd1{xbwd_args(a), t, nch, off}, W,
                                            H, nimg * W, ig, a.Lp, a.p0, nch * a.K, t, lb, nb);
  } else {
    xlse_pass_body<true, XEpiBwd2, kXWaves>(a.xt + off, a.sgt + off, XEpiBwd2{xbwd_args(a), t, nch}, W, H,
                                            nimg * W, ig, a.Lp, a.p0, nch * a.K, t, lb, nb);
  }
}

template <bool FUSED>
__global__ __launch_bounds__(kThreads) void k_xdot_chunk_group(GroupArgs g, int t) {
  int lb, nb;
  const GroupMember& gm = g.m[group_member(g, GK_DOT, blockIdx.x, lb, nb)];
  if (*gm.lmax < t) return;
  const SolveArgs& a = gm.a;
  xdot_chunk_body<FUSED>(a, t, lb);
}

__global__ void k_xb1_group(GroupArgs g, int t) {
  int lb, nb;
  const GroupMember& gm = g.m[group_member(g, GK_XB1, blockIdx.x, lb, nb)];
  if (*gm.lmax < t) return;
  const SolveArgs& a = gm.a;
  xb1_body(a, t, a.weightsOnly ? 1 : 2, lb, nb);
}

__global__ void k_xdot_fin_group(GroupArgs g, int t) {
  int lb, nb;
  const GroupMember& gm = g.m[group_member(g, GK_DOTFIN, blockIdx.x, lb, nb)];
  if (*gm.lmax < t) return;
  const SolveArgs& a = gm.a;
  xdot_fin_body(a, t, lb, nb);
}

int xlse_group_set_attr(int lds) {
  static std::atomic<int> done{0};  // groups may be created from several host threads
  if (lds <= done.load(std::memory_order_relaxed)) return WDL_OK;
  const void* fns[] = {(const void*)k_xlse_group_fwd<0>, (const void*)k_xlse_group_fwd<1>,
                       (const void*)k_xlse_group_fwd<2>, (const void*)k_xlse_group_fwd<3>,
                       (const void*)k_xlse_group_bwd<0>, (const void*)k_xlse_group_bwd<1>,
                       (const void*)k_xlse_group_bwd<2>, (const void*)k_xlse_group_bwd<3>};
  for (const void* f : fns)
    WDL_HIP(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, lds));
  int cur = done.load(std::memory_order_relaxed);
  while (cur < lds && !done.compare_exchange_weak(cur, lds, std::memory_order_relaxed)) {
  }
  return WDL_OK;
}

void group_blocks(const std::vector<SolveArgs>& ms, int lds, int cus,
                  std::vector<int> (&counts)[GK_COUNT]) {
  const size_t nm = ms.size();
  for (auto& c : counts) c.assign(nm, 0);
  auto ew = [](long long n) { return (int)std::min<long long>((n + 255) / 256, 8192); };
  // the passes: each member's lines in groups of kXWaves; when the members
  // need more workgroups than twice what the device holds at once (the cap a
  // single-solver pass takes), the cap is shared out by each member's work
  // (lines x length), at least one workgroup each
  const int perCU = std::max(1, std::min(160 * 1024 / lds, 8));
  const long long cap = 2LL * (cus > 0 ? cus : 256) * perCU;
  auto passes = [&](int kind, bool bwd, bool alongX) {
    std::vector<long long> need(nm), work(nm);
    long long tneed = 0, twork = 0;
    for (size_t m = 0; m < nm; ++m) {
      const SolveArgs& a = ms[m];
      const int nch = bwd && !a.weightsOnly ? 2 : 1;
      const long long nimg = (long long)a.np * nch * a.K;
      const int rows = alongX ? a.G.H : a.G.W, len = alongX ? a.G.W : a.G.H;
      const long long lines = nimg * rows;
      need[m] = (lines + kXWaves - 1) / kXWaves;
      work[m] = lines * len;
      tneed += need[m];
      twork += work[m];
    }
    for (size_t m = 0; m < nm; ++m) {
      long long c = need[m];
      if (tneed > cap)
        c = std::min(need[m], std::max(1LL, (long long)((double)cap * work[m] / twork)));
      counts[kind][m] = (int)c;
    }
  };
  passes(GK_FA, false, true);
  passes(GK_FB, false, false);
  passes(GK_BA, true, true);
  passes(GK_BB, true, false);
  for (size_t m = 0; m < nm; ++m) {
    const SolveArgs& a = ms[m];
    const long long N = a.N, KN = (long long)a.K * N;
    counts[GK_XF1][m] = ew((long long)a.np * N);
    counts[GK_FINAL][m] = kLossBlocks * a.np;
    counts[GK_DOT][m] = (int)((long long)a.np * a.K * xdot_chunks(N));
    counts[GK_XB1][m] = a.weightsOnly ? 0 : ew((long long)a.np * 2 * KN);
    counts[GK_DOTFIN][m] = ew((long long)a.np * a.K);
    counts[GK_PROJ][m] = (a.P + 63) / 64;
    counts[GK_WARM][m] = (int)std::min<long long>(64, (KN + 255) / 256) * a.P;
    counts[GK_RGD][m] = a.weightsOnly ? 0 : (int)std::min<long long>(1024, (KN + 255) / 256);
  }
}

int exact_group_forward(const GroupArgs& g, const GroupShape& sh, hipStream_t st, int s) {
  const dim3 pb(kXWaves * 64 * sh.wpl);
  hipLaunchKernelGGL(k_xf1_group, dim3(sh.grid[GK_XF1]), dim3(256), 0, st, g, s);
  hipLaunchKernelGGL(k_xlse_group_fwd<0>, dim3(sh.grid[GK_FA]), pb, sh.lds, st, g, s);
  hipLaunchKernelGGL(k_xlse_group_fwd<1>, dim3(sh.grid[GK_FB]), pb, sh.lds, st, g, s);
  hipLaunchKernelGGL(k_xlse_group_fwd<2>, dim3(sh.grid[GK_FA]), pb, sh.lds, st, g, s);
  hipLaunchKernelGGL(k_xlse_group_fwd<3>, dim3(sh.grid[GK_FB]), pb, sh.lds, st, g, s);
  WDL_HIP(hipGetLastError());
  return WDL_OK;
}

int exact_group_backward(const GroupArgs& g, const GroupShape& sh, hipStream_t st, int t) {
  const dim3 pb(kXWaves * 64 * sh.wpl);
  if (sh.weightsOnly) {
    hipLaunchKernelGGL(k_xdot_chunk_group<true>, dim3(sh.grid[GK_DOT]), dim3(kThreads), 0, st, g,
                       t);
  } else {
    hipLaunchKernelGGL(k_xdot_chunk_group<false>, dim3(sh.grid[GK_DOT]), dim3(kThreads), 0, st, g,
                       t);
    hipLaunchKernelGGL(k_xb1_group, dim3(sh.grid[GK_XB1]), dim3(256), 0, st, g, t);
  }
  hipLaunchKernelGGL(k_xdot_fin_group, dim3(sh.grid[GK_DOTFIN]), dim3(256), 0, st, g, t);
  hipLaunchKernelGGL(k_xlse_group_bwd<0>, dim3(sh.grid[GK_BA]), pb, sh.lds, st, g, t);
  hipLaunchKernelGGL(k_xlse_group_bwd<1>, dim3(sh.grid[GK_BB]), pb, sh.lds, st, g, t);
  if (t >= 2) {
    hipLaunchKernelGGL(k_xlse_group_bwd<2>, dim3(sh.grid[GK_BA]), pb, sh.lds, st, g, t);
    hipLaunchKernelGGL(k_xlse_group_bwd<3>, dim3(sh.grid[GK_BB]), pb, sh.lds, st, g, t);
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
  XGRID;
  const long long tot = (long long)a.np * a.N, off = (long long)a.p0 * a.N;
  XLOOP(tot) a.wlb[off + e] = 0.0;
}

// log a = log q - conv (second = false) or log b = log o - conv (second = true)
__global__ void k_wl_update(SolveArgs a, int second) {
  XGRID;
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
#undef XGRID

}  // namespace wdl
