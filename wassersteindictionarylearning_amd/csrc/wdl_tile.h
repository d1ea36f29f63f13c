// Tiled separable Gaussian convolution for images larger than the LDS holds.
//
// A workgroup computes one 64 x 64 output tile.  It loads the tile plus a halo
// of HB pixels on every side (HB >= D, a multiple of 8) into LDS, runs pass 1
// (along x) over all R = 64 + 2 HB haloed rows for the 64 inner columns,
// writes the result transposed, and runs pass 2 (along y) over the R haloed
// positions for the 64 inner rows.  Factors beyond D are exactly 0 in FP32,
// so dropping everything outside the halo changes no bit of the result
// compared with the dense product.  The log-domain form applies the shifted
// scheme of wdl_logconv.h to the same shapes.
#pragma once

#include "wdl_logconv.h"

namespace wdl {

constexpr int kTile = 64;
constexpr int kHaloMax = 32;  // R = 64 + 2 HB <= 128 = kMaxSide

struct TileGeom {
  int W, H;    // image
  int HB;      // halo
  int R;       // haloed region side
  int S;       // LDS row stride (region rows and transposed rows)
  int D;       // band half-width
  int nx, ny;  // tiles per row, per column
};

__host__ __device__ inline int tile_lds_bytes(const TileGeom& t) {
  const int img = t.R * t.S;
  return (kTabFloats + img) * (int)sizeof(float);
}

// Visits the haloed region of tile (tx, ty): region position (r, k) <-> image
// pixel (ty*64 - HB + r, tx*64 - HB + k); outside the image `ok` is false.
// One group per thread at a time: the log-domain loads carry FP64 values, and
// fewer of them in flight keep the kernels at three workgroups per CU.
constexpr int kUnrollTile = 1;
template <int VW, class Load, class Apply>
__device__ inline void tile_region_pass(const TileGeom& T, int tx, int ty, Load&& load,
                                        Apply&& apply) {
  const int q = T.R / VW;
  const int total = T.R * q;
  const int y0 = ty * kTile - T.HB, x0 = tx * kTile - T.HB;
  typedef decltype(load(0)) V;
  for (int base = threadIdx.x; base < total; base += kThreads * kUnrollTile) {
    V v[kUnrollTile];
    int li[kUnrollTile], gi[kUnrollTile];
    bool ok[kUnrollTile];
#pragma unroll
    for (int u = 0; u < kUnrollTile; ++u) {
      const int e = base + u * kThreads;
      const int r = e / q, k = (e - r * q) * VW;
      const int y = y0 + r, x = x0 + k;
      li[u] = r * T.S + k;
      ok[u] = e < total && y >= 0 && y < T.H && x >= 0 && x < T.W;
      gi[u] = ok[u] ? y * T.W + x : 0;
      if (ok[u]) v[u] = load(gi[u]);
    }
#pragma unroll
    for (int u = 0; u < kUnrollTile; ++u)
      if (base + u * kThreads < total) apply(li[u], gi[u], ok[u], v[u]);
  }
}

// Visits the 64 x 64 inner tile as laid out by tile_conv's result (stride S).
template <int VW, class Load, class Apply>
__device__ inline void tile_inner_pass(const TileGeom& T, int tx, int ty, Load&& load,
                                       Apply&& apply) {
  const int q = kTile / VW;
  const int total = kTile * q;
  typedef decltype(load(0)) V;
  for (int base = threadIdx.x; base < total; base += kThreads * kUnroll) {
    V v[kUnroll];
    int li[kUnroll], gi[kUnroll];
    bool ok[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      const int e = base + u * kThreads;
      const int i = e / q, j = (e - i * q) * VW;
      const int y = ty * kTile + i, x = tx * kTile + j;
      li[u] = i * T.S + j;
      ok[u] = e < total && y < T.H && x < T.W;
      gi[u] = ok[u] ? y * T.W + x : 0;
      if (ok[u]) v[u] = load(gi[u]);
    }
#pragma unroll
    for (int u = 0; u < kUnroll; ++u)
      if (base + u * kThreads < total) apply(li[u], gi[u], ok[u], v[u]);
  }
}

// buf holds the haloed input region (R x R, stride S, zero outside the image);
// on return it holds the 64 x 64 output tile at stride S.
__device__ inline void tile_conv(float* buf, const float* tabs, const TileGeom& T) {
  floatx4 acc[kMaxUnits][4];
  __syncthreads();
  rows_gemm(buf, T.S, T.R, T.R, kTile, T.HB, tabs, T.D, acc);     // region rows, inner columns
  __syncthreads();
  store_transposed(buf, T.S, T.R, kTile, acc);                     // 64 lines of R
  __syncthreads();
  rows_gemm(buf, T.S, kTile, T.R, kTile, T.HB, tabs, T.D, acc);   // inner rows
  __syncthreads();
  store_transposed(buf, T.S, kTile, kTile, acc);                   // 64 x 64, row-major
  __syncthreads();
}

// Whether region LDS index li lies in the inner 64 x 64 tile (whole VW
// groups: HB and the group start are multiples of 4).
__device__ inline bool tile_inner(const TileGeom& T, int li) {
  const int r = li / T.S, k = li - r * T.S;
  return (unsigned)(r - T.HB) < (unsigned)kTile && (unsigned)(k - T.HB) < (unsigned)kTile;
}

// Log-domain tile convolution (the shifted scheme of log_conv_lds on the
// tile's shapes).  The input of region pixel group li / gi is
// make(li, gi, v, second) -> SLog<VW> with v = load(gi).  With groups of 4
// pixels make is evaluated once per group (second = true); single pixels
// evaluate it twice, for the row maxima (second = false) and for the shifted
// values (second = true), so side outputs belong in the second = true call.
// make must not read buf.  On return buf holds S2 (64 x 64 at stride S) and
// L.mcol the shifts of the inner columns.
template <int VW, class Load, class Make>
__device__ inline void tile_log_conv(float* buf, const float* tabs, const LogLds& L,
                                     const TileGeom& T, int tx, int ty, unsigned* events,
                                     Load&& load, Make&& make) {
  typedef typename Vec<VW>::T V;
  if constexpr (VW == 4) {
    // One evaluation of the region: a half-wave (32 lanes) per region row, one
    // group of 4 pixels per lane (R / 4 <= 32 groups), so the row's maximum
    // and finite minimum are half-wave DPP reductions and each value is
    // shifted by its row's maximum as soon as it is made -- one round of
    // loads and one make() per pixel.  R is a multiple of 16, so every wave
    // runs the same R / 16 rows' worth of steps with all lanes active.
    const int q = T.R / 4;
    const int lane = threadIdx.x & 63, l32 = lane & 31;
    const int y0 = ty * kTile - T.HB, x0 = tx * kTile - T.HB;
    for (int r = threadIdx.x >> 5; r < T.R; r += kThreads / 32) {
      const int y = y0 + r, x = x0 + l32 * 4;
      const int li = r * T.S + l32 * 4;
      const bool ok = l32 < q && y >= 0 && y < T.H && x >= 0 && x < T.W;
      const int gi = ok ? y * T.W + x : 0;
      SLog<VW> z{};
      float mx = -INFINITY, mn = INFINITY;
      if (ok) {
        z = make(li, gi, load(gi), true);
        const V xf = vf(z.x);
        mx = vmax_all(xf);
        mn = vmin_fin(xf);
      }
      const float hx = half_scan(mx, MaxOp());
      const float m = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(hx), 31));
      const float m2 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(hx), 63));
      const float rm = (lane & 32) ? m2 : m;
      // the range check: a finite value more than kLogRange under the row's
      // maximum (one row per half of the ballot)
      const unsigned long long bad = __ballot(rm - mn > kLogRange);
      if (l32 < q) {
        V e = vsplat<VW>(0.f);
        if (ok && rm > -INFINITY) e = z.s * vexp_shift(z.x, rm);
        vst<VW>(buf + li, e);
      }
      if (l32 == 0) {
        L.mrow[r] = fkey(rm);  // rows outside the image stay at -inf
        if (events && ((bad >> (lane & 32)) & 0xffffffffull)) atomicAdd(events, 1u);
      }
    }
  } else {
    for (int r = threadIdx.x; r < T.R; r += kThreads) {
      L.mrow[r] = fkey(-INFINITY);
      L.rmin[r] = fkey(INFINITY);
    }
    __syncthreads();
  }
  if constexpr (VW != 4) {
    // single pixels (widths not a multiple of 4): row keys by LDS atomics
    // from a first evaluation, then the shifted values from a second
    tile_region_pass<VW>(T, tx, ty, load, [&](int li, int gi, bool ok, const auto& v) {
      if (ok) {
        const SLog<VW> z = make(li, gi, v, false);
        const V xf = vf(z.x);
        const int r = li / T.S;
        atomicMax(L.mrow + r, fkey(vmax_all(xf)));
        atomicMin(L.rmin + r, fkey(vmin_fin(xf)));
      }
    });
    __syncthreads();
    check_range(L.mrow, L.rmin, T.R, events);  // region rows outside the image stay at -inf
    tile_region_pass<VW>(T, tx, ty, load, [&](int li, int gi, bool ok, const auto& v) {
      V e = vsplat<VW>(0.f);
      if (ok) {
        const SLog<VW> z = make(li, gi, v, true);
        const float m = funkey(L.mrow[li / T.S]);
        if (m > -INFINITY) e = z.s * vexpf(vf(z.x - (double)m));
      }
      vst<VW>(buf + li, e);
    });
  }
  floatx4 acc[kMaxUnits][4];
  __syncthreads();
  rows_gemm(buf, T.S, T.R, T.R, kTile, T.HB, tabs, T.D, acc);
  __syncthreads();
  store_transposed(buf, T.S, T.R, kTile, acc);  // line c (inner column), position r (region row)
  __syncthreads();
  const int ncol = min(kTile, T.W - tx * kTile);  // inner columns inside the image
  // Per inner column (one wave each, R <= 128 positions in two per lane, kept
  // in registers): the maximum of T = M_r + log|S1| by a DPP scan, the range
  // check by a ballot, then E2 = S1 exp(M_r - M'_c) written back in place.
  const int lane = threadIdx.x & 63, wave = wave_id();
  const bool has1 = lane + 64 < T.R;
  const float mr0 = funkey(L.mrow[lane]);
  const float mr1 = has1 ? funkey(L.mrow[lane + 64]) : -INFINITY;
  for (int c = wave; c < kTile; c += kWaves) {
    float* col = buf + c * T.S;
    const float v0 = col[lane], v1 = has1 ? col[lane + 64] : 0.f;
    float t0 = -INFINITY, t1 = -INFINITY;
    if (c < ncol) {
      if (mr0 > -INFINITY) t0 = mr0 + fast_logf(fabsf(v0));
      if (mr1 > -INFINITY) t1 = mr1 + fast_logf(fabsf(v1));
    }
    const float m = wave_max(fmaxf(t0, t1));
    const float lo = fminf(vmin_fin(t0), vmin_fin(t1));
    const unsigned long long bad = __ballot(m - lo > kLogRange);
    if (lane == 0) {
      L.mcol[c] = fkey(m);
      if (events && bad) atomicAdd(events, 1u);
    }
    // a factor beyond FP32 range only occurs for |S1| < e^-80 and goes
    // through logs
    auto e2 = [&](float v, float mr) {
      if (!(c < ncol && m > -INFINITY && mr > -INFINITY)) return 0.f;
      const float d = mr - m;
      return d < 80.f ? v * vexp_shift((double)mr, m)
                      : vsign(v) * expf(logf(fabsf(v)) + d);
    };
    col[lane] = e2(v0, mr0);
    if (has1) col[lane + 64] = e2(v1, mr1);
  }
  __syncthreads();
  rows_gemm(buf, T.S, kTile, T.R, kTile, T.HB, tabs, T.D, acc);
  __syncthreads();
  store_transposed(buf, T.S, kTile, kTile, acc);
  __syncthreads();
}

// log|R| and sign at inner-tile LDS index li (row i = li / S).
template <int VW>
__device__ inline SLog<VW> tile_log_result(const float* buf, const LogLds& L, const TileGeom& T,
                                           int li) {
  typedef typename Vec<VW>::T V;
  const V r = vld<VW>(buf + li);
  const int j = li - (li / T.S) * T.S;
  SLog<VW> z;
  z.x = vd(mcol_at(L.mcol, j, V())) + vhlog(vabs(r));
  z.s = vsign(r);
  return z;
}

}  // namespace wdl
