// Log-domain separable Gaussian convolution on MFMA, image resident in LDS.
//
// Input: per pixel a log-magnitude X and a sign s (+-1).  Output: per pixel
// log|R| and sign(R) of R = K_H . (s exp X) . K_W.  The exact per-output max
// of the reference (kernels.h:1082-1091) is replaced by shifts that make both
// passes plain GEMMs:
//   pass 1: M_r = max_k X[r][k];  E = s exp(X - M_r);  S1 = E . K_W       (MFMA)
//   pass 2: T = M_r + log|S1|;  M'_c = max_r T[r][c];
//           E2 = sign(S1) exp(T - M'_c);  S2 = K_H . E2                  (MFMA)
//   result: log|R| = M'_c + log|S2|,  sign(R) = sign(S2).
// Every shifted term is <= 1, so nothing overflows.  A term is lost to FP32
// underflow only if it lies more than ~87 nats under its line's maximum;
// such terms are negligible for every output as long as each line's spread
// of finite values stays below kLogRange (the diagonal factor is 1, the
// smallest kept factor exp(-103)), which is checked on every line: lines
// beyond it are counted in *range_events for the host to report.
#pragma once

#include "wdl_conv_core.h"

namespace wdl {

constexpr float kLogRange = 80.f;

// float <-> int key, monotonic, for LDS atomicMax / atomicMin on floats.
__device__ inline int fkey(float f) {
  const int i = __float_as_int(f);
  return i >= 0 ? i : i ^ 0x7FFFFFFF;
}
__device__ inline float funkey(int k) { return __int_as_float(k >= 0 ? k : k ^ 0x7FFFFFFF); }

struct LogLds {
  int* mrow;  // kMaxSide keys: row max, then row min during the range check
  int* mcol;  // kMaxSide keys
  int* rmin;  // kMaxSide keys
};

// Log values are carried in FP64 inside a kernel (histories are stored in FP32):
// every shift difference X - M is then exact before the FP32 exp.
typedef double doublex4 __attribute__((ext_vector_type(4)));
template <int VW> struct VecD;
template <> struct VecD<1> { typedef double T; };
template <> struct VecD<4> { typedef doublex4 T; };
__device__ inline double vd(float x) { return (double)x; }
__device__ inline doublex4 vd(floatx4 x) { return __builtin_convertvector(x, doublex4); }
__device__ inline float vf(double x) { return (float)x; }
__device__ inline floatx4 vf(doublex4 x) { return __builtin_convertvector(x, floatx4); }

// log of a linear FP32 value as an FP64 log value: the binary exponent goes
// in exactly, only log of the mantissa (in [0.5, 1), |log| < 0.7) is
// rounded, so the absolute error stays ~4e-8 however large |log x| is.
// exp(x - m) for an FP64 log value x and an FP32 shift m: the difference and
// its scaling to log2 units in FP64, rounded once to FP32, then the hardware
// exp2 -- the same single rounding of the argument as expf(float(x - m)) has,
// without expf's range reduction.
__device__ inline float vexp_shift(double x, float m) {
  return __builtin_amdgcn_exp2f((float)((x - (double)m) * 1.4426950408889634));
}
__device__ inline floatx4 vexp_shift(doublex4 x, float m) {
  return floatx4{vexp_shift(x.x, m), vexp_shift(x.y, m), vexp_shift(x.z, m), vexp_shift(x.w, m)};
}

// exp(d) of an FP64 exponent, by the same route as vexp_shift.
__device__ inline float vexp_d(double d) {
  return __builtin_amdgcn_exp2f((float)(d * 1.4426950408889634));
}
__device__ inline floatx4 vexp_d(doublex4 d) {
  return floatx4{vexp_d(d.x), vexp_d(d.y), vexp_d(d.z), vexp_d(d.w)};
}

// The mantissa's log2 on the hardware log (within an ulp of a value in
// [-1, 0), i.e. ~6e-8 absolute), the exponent added exactly in FP64.
__device__ inline double hlog(float x) {
  int e;
  const float m = frexpf(x, &e);
  return ((double)e + (double)__builtin_amdgcn_logf(m)) * 0.69314718055994530942;
}
__device__ inline double vhlog(float x) { return hlog(x); }
__device__ inline doublex4 vhlog(floatx4 x) {
  return doublex4{hlog(x.x), hlog(x.y), hlog(x.z), hlog(x.w)};
}
// hlog for a make's first evaluation (line maxima only: fast_logf), the
// accurate one for its second (the values that are convolved).
__device__ inline double vhlog_sel(float x, bool second) {
  return second ? hlog(x) : (double)fast_logf(x);
}
__device__ inline doublex4 vhlog_sel(floatx4 x, bool second) {
  if (second) return vhlog(x);
  return doublex4{fast_logf(x.x), fast_logf(x.y), fast_logf(x.z), fast_logf(x.w)};
}

template <int VW>
__device__ inline typename VecD<VW>::T vldd(const double* p) {
  return *reinterpret_cast<const typename VecD<VW>::T*>(p);
}
template <int VW>
__device__ inline void vstd(double* p, typename VecD<VW>::T v) {
  *reinterpret_cast<typename VecD<VW>::T*>(p) = v;
}

template <int VW> struct SLog {
  typename VecD<VW>::T x;  // log magnitude
  typename Vec<VW>::T s;   // +1 / -1
};

__device__ inline float vmax_all(float x) { return x; }
__device__ inline float vmax_all(floatx4 x) { return fmaxf(fmaxf(x.x, x.y), fmaxf(x.z, x.w)); }
__device__ inline float vmin_fin(float x) { return isinf(x) ? INFINITY : x; }
__device__ inline float vmin_fin(floatx4 x) {
  return fminf(fminf(vmin_fin(x.x), vmin_fin(x.y)), fminf(vmin_fin(x.z), vmin_fin(x.w)));
}
__device__ inline float vsign(float x) { return x >= 0.f ? 1.f : -1.f; }
__device__ inline floatx4 vsign(floatx4 x) {
  return floatx4{vsign(x.x), vsign(x.y), vsign(x.z), vsign(x.w)};
}
__device__ inline float vabs(float x) { return fabsf(x); }
__device__ inline floatx4 vabs(floatx4 x) {
  return floatx4{fabsf(x.x), fabsf(x.y), fabsf(x.z), fabsf(x.w)};
}

// Like px_pass, also handing the row index to apply.
template <int VW, class Load, class Apply>
__device__ inline void px_pass_r(const ConvGeom& G, Load&& load, Apply&& apply) {
  const int q = G.Wp / VW;
  const int total = G.Hp * q;
  typedef decltype(load(0)) V;
  // the incremental group walk of px_pass
  const int drow = kThreads / q, dcol = kThreads - drow * q;
  int r = threadIdx.x / q, col = threadIdx.x - r * q;
  int lrow = r * G.S1 + col * VW, grow = r * G.W + col * VW;
  for (int base = threadIdx.x; base < total; base += kThreads * kUnroll) {
    V v[kUnroll];
    int li[kUnroll], gi[kUnroll], row[kUnroll];
    bool ok[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      row[u] = r;
      li[u] = lrow;
      gi[u] = grow;
      ok[u] = r < G.H && col * VW < G.W;
      if (ok[u]) v[u] = load(gi[u]);
      r += drow;
      col += dcol;
      lrow += drow * G.S1 + dcol * VW;
      grow += drow * G.W + dcol * VW;
      if (col >= q) {
        col -= q;
        ++r;
        lrow += G.S1 - G.Wp;
        grow += G.W - G.Wp;
      }
    }
#pragma unroll
    for (int u = 0; u < kUnroll; ++u)
      if (base + u * kThreads < total) apply(li[u], gi[u], row[u], ok[u], v[u]);
  }
}

// Checks every line's spread (max - finite min) and counts lines beyond kLogRange.
__device__ inline void check_range(const int* mx, const int* mn, int lines, unsigned* events) {
  int bad = 0;
  for (int r = threadIdx.x; r < lines; r += kThreads) {
    const float hi = funkey(mx[r]), lo = funkey(mn[r]);
    if (isfinite(hi) && isfinite(lo) && hi - lo > kLogRange) ++bad;
  }
  if (bad && events) atomicAdd(events, (unsigned)bad);
}

// The conv input is produced by load(gi) (global reads, batched) and
// make(li, gi, row, v, second) -> SLog<VW>.  make may read LDS at li and is
// evaluated twice per pixel group, first for the row maxima (second = false),
// then for the shifted values (second = true); side outputs belong in the
// second call.
// On return buf holds S2 (row-major, stride S1) and L.mcol the column keys.
template <int VW, class Load, class Make>
__device__ inline void log_conv_lds(float* buf, const float* tabs, const LogLds& L,
                                    const ConvGeom& G, unsigned* events, Load&& load, Make&& make) {
  typedef typename Vec<VW>::T V;
  for (int r = threadIdx.x; r < kMaxSide; r += kThreads) {
    L.mrow[r] = fkey(-INFINITY);
    L.rmin[r] = fkey(INFINITY);
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = wave_id();
  // row maxima (and finite minima for the range check)
  if constexpr (VW == 4) {
    // Two rows per wave step, 32 lanes each (a row has at most 32 float4
    // groups), reduced with DPP scans: no contended LDS atomics.  make only
    // reads LDS here.
    const int q = G.Wp / 4, half = lane >> 5, gcol = (lane & 31) * 4;
    for (int r0 = 2 * wave; r0 < G.Hp; r0 += 2 * kWaves) {
      const int row = r0 + half;
      float mx = -INFINITY, mn = INFINITY;
      if (row < G.H && (lane & 31) < q && gcol < G.W) {
        const int li = row * G.S1 + gcol, gi = row * G.W + gcol;
        const V xf = vf(make(li, gi, row, load(gi), false).x);
        mx = vmax_all(xf);
        mn = vmin_fin(xf);
      }
      mx = half_scan(mx, MaxOp());  // lanes 31 and 63 hold their half's
      mn = half_scan(mn, MinOp());
      if ((lane & 31) == 31 && row < G.H) {
        L.mrow[row] = fkey(mx);
        L.rmin[row] = fkey(mn);
      }
    }
  } else {
    px_pass_r<VW>(G, load, [&](int li, int gi, int row, bool ok, const auto& v) {
      if (ok) {
        const SLog<VW> z = make(li, gi, row, v, false);
        const V xf = vf(z.x);
        atomicMax(L.mrow + row, fkey(vmax_all(xf)));
        atomicMin(L.rmin + row, fkey(vmin_fin(xf)));
      }
    });
  }
  __syncthreads();
  check_range(L.mrow, L.rmin, G.H, events);
  // shifted, signed values
  px_pass_r<VW>(G, load, [&](int li, int gi, int row, bool ok, const auto& v) {
    V e = vsplat<VW>(0.f);
    if (ok) {
      const SLog<VW> z = make(li, gi, row, v, true);
      const float m = funkey(L.mrow[row]);
      if (m > -INFINITY) e = z.s * vexpf(vf(z.x - (double)m));
    }
    vst<VW>(buf + li, e);
  });
  floatx4 acc[kMaxUnits][4];
  __syncthreads();
  rows_gemm(buf, G.S1, G.Hp, G.Wp, tabs, G.D, acc);  // S1 = E . K_W
  __syncthreads();
  store_transposed(buf, G.S2, G.Hp, G.Wp, acc);       // buf = S1^T (Wp lines of Hp)
  __syncthreads();
  // Column maxima (and finite minima) of T = M_r + log|S1|: one wave per
  // line c (image column), reduced with DPP scans (no LDS atomics).
  for (int c = wave; c < kMaxSide; c += kWaves) {
    float mx = -INFINITY, mn = INFINITY;
    if (c < G.W)
      for (int r = lane; r < G.H; r += 64) {
        const float t = funkey(L.mrow[r]) + fast_logf(fabsf(buf[c * G.S2 + r]));
        mx = fmaxf(mx, t);
        mn = fminf(mn, vmin_fin(t));
      }
    mx = wave_max(mx);
    mn = wave_min(mn);
    if (lane == 0) {
      L.mcol[c] = fkey(mx);
      L.rmin[c] = fkey(mn);
    }
  }
  __syncthreads();
  check_range(L.mcol, L.rmin, G.W, events);
  for (int c = wave; c < G.Wp; c += kWaves) {
    const float m = c < G.W ? funkey(L.mcol[c]) : -INFINITY;
    for (int r = lane; r < G.Hp; r += 64) {
      float* p = buf + c * G.S2 + r;
      float v = 0.f;
      if (r < G.H && m > -INFINITY) {
        // E2 = S1 exp(M_r - M'_c), the shift difference in FP64; a factor
        // beyond FP32 range only occurs for |S1| < e^-80 and goes through logs.
        const float d = (float)((double)funkey(L.mrow[r]) - (double)m);
        v = d < 80.f ? *p * expf(d) : vsign(*p) * expf(logf(fabsf(*p)) + d);
      }
      *p = v;
    }
  }
  __syncthreads();
  rows_gemm(buf, G.S2, G.Wp, G.Hp, tabs, G.D, acc);  // S2^T = E2^T . K_H
  __syncthreads();
  store_transposed(buf, G.S1, G.Wp, G.Hp, acc);       // buf = S2 (Hp x Wp)
  __syncthreads();
}

// Column shifts of pass 2 for the VW columns starting at col.
__device__ inline float mcol_at(const int* mcol, int col, float) { return funkey(mcol[col]); }
__device__ inline floatx4 mcol_at(const int* mcol, int col, floatx4) {
  return floatx4{funkey(mcol[col]), funkey(mcol[col + 1]), funkey(mcol[col + 2]),
                 funkey(mcol[col + 3])};
}

// log|R| of the pixel group at li (after log_conv_lds), and its sign.
template <int VW>
__device__ inline SLog<VW> log_result(const float* buf, const LogLds& L, const ConvGeom& G, int li,
                                      int row) {
  typedef typename Vec<VW>::T V;
  const V r = vld<VW>(buf + li);
  SLog<VW> z;
  z.x = vd(mcol_at(L.mcol, li - row * G.S1, V())) + vhlog(vabs(r));
  z.s = vsign(r);
  return z;
}

constexpr int kLogKeysBytes = 3 * kMaxSide * (int)sizeof(int);

}  // namespace wdl
