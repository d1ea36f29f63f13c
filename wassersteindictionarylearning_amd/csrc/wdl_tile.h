// Tiled separable Gaussian convolution for images larger than the LDS holds.
//
// A workgroup computes one 64 x 64 output tile.  It loads the tile plus a halo
// of HB pixels on every side (HB >= D, a multiple of 16) into LDS, runs pass 1
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

// tile_region_pass with the group index e = r * (R / VW) + k / VW handed to
// apply(e, li, gi, valid, value) as well.
template <int VW, class Load, class Apply>
__device__ inline void tile_region_pass_e(const TileGeom& T, int tx, int ty, Load&& load,
                                          Apply&& apply) {
  tile_region_pass<VW>(T, tx, ty, load, [&](int li, int gi, bool ok, const auto& v) {
    const int q = T.R / VW;
    const int r = li / T.S;
    apply(r * q + (li - r * T.S) / VW, li, gi, ok, v);
  });
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
// make(li, gi, v, second) -> SLog<VW> with v = load(gi); make is evaluated
// twice, for the row maxima (second = false) and for the shifted values
// (second = true), so side outputs belong in the second call.  make must not
// read buf: the first evaluation stages its row keys there.  On return buf
// holds S2 (64 x 64 at stride S) and L.mcol the shifts of the inner columns.
template <int VW, class Load, class Make>
__device__ inline void tile_log_conv(float* buf, const float* tabs, const LogLds& L,
                                     const TileGeom& T, int tx, int ty, unsigned* events,
                                     Load&& load, Make&& make) {
  typedef typename Vec<VW>::T V;
  for (int r = threadIdx.x; r < kMaxSide; r += kThreads) {
    L.mrow[r] = fkey(-INFINITY);
    L.rmin[r] = fkey(INFINITY);
  }
  __syncthreads();
  if constexpr (VW == 4) {
    // Row maxima (and finite minima) without contended LDS atomics: each
    // group's keys go to buf (free until the shifted values are written),
    // then one thread per region row reduces its R / 4 groups.
    const int q = T.R / 4, ng = T.R * q;
    int* gmax = reinterpret_cast<int*>(buf);
    int* gmin = gmax + ng;
    tile_region_pass_e<VW>(T, tx, ty, load, [&](int e, int li, int gi, bool ok, const auto& v) {
      int kx = fkey(-INFINITY), kn = fkey(INFINITY);
      if (ok) {
        const V xf = vf(make(li, gi, v, false).x);
        kx = fkey(vmax_all(xf));
        kn = fkey(vmin_fin(xf));
      }
      gmax[e] = kx;
      gmin[e] = kn;
    });
    __syncthreads();
    for (int r = threadIdx.x; r < T.R; r += kThreads) {
      int kx = fkey(-INFINITY), kn = fkey(INFINITY);
      for (int g = 0; g < q; ++g) {
        kx = max(kx, gmax[r * q + g]);
        kn = min(kn, gmin[r * q + g]);
      }
      L.mrow[r] = kx;
      L.rmin[r] = kn;
    }
  } else {
    tile_region_pass<VW>(T, tx, ty, load, [&](int li, int gi, bool ok, const auto& v) {
      if (ok) {
        const SLog<VW> z = make(li, gi, v, false);
        const V xf = vf(z.x);
        const int r = li / T.S;
        atomicMax(L.mrow + r, fkey(vmax_all(xf)));
        atomicMin(L.rmin + r, fkey(vmin_fin(xf)));
      }
    });
  }
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
  floatx4 acc[kMaxUnits][4];
  __syncthreads();
  rows_gemm(buf, T.S, T.R, T.R, kTile, T.HB, tabs, T.D, acc);
  __syncthreads();
  store_transposed(buf, T.S, T.R, kTile, acc);  // line c (inner column), position r (region row)
  __syncthreads();
  const int ncol = min(kTile, T.W - tx * kTile);  // inner columns inside the image
  // Column maxima (and finite minima) of T = M_r + log|S1|: one wave per
  // inner column, reduced with DPP scans (no LDS atomics, no index division).
  const int lane = threadIdx.x & 63, wave = wave_id();
  for (int c = wave; c < kTile; c += kWaves) {
    float mx = -INFINITY, mn = INFINITY;
    if (c < ncol)
      for (int r = lane; r < T.R; r += 64) {
        const float m = funkey(L.mrow[r]);
        if (m > -INFINITY) {
          const float t = m + fast_logf(fabsf(buf[c * T.S + r]));
          mx = fmaxf(mx, t);
          mn = fminf(mn, vmin_fin(t));
        }
      }
    mx = wave_max(mx);
    mn = wave_min(mn);
    if (lane == 0) {
      L.mcol[c] = fkey(mx);
      L.rmin[c] = fkey(mn);
    }
  }
  __syncthreads();
  check_range(L.mcol, L.rmin, kTile, events);
  for (int c = wave; c < kTile; c += kWaves) {
    const float m = funkey(L.mcol[c]);
    for (int r = lane; r < T.R; r += 64) {
      float* p = buf + c * T.S + r;
      float v = 0.f;
      const float mr = funkey(L.mrow[r]);
      if (c < ncol && m > -INFINITY && mr > -INFINITY) {
        // E2 = S1 exp(M_r - M'_c), the shift difference in FP64; a factor
        // beyond FP32 range only occurs for |S1| < e^-80 and goes through logs.
        const float d = (float)((double)mr - (double)m);
        v = d < 80.f ? *p * expf(d) : vsign(*p) * expf(logf(fabsf(*p)) + d);
      }
      *p = v;
    }
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
