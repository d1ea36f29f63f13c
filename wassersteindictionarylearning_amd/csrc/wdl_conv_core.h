// Device-side building blocks for the separable Gaussian convolution on gfx950.
//
// One workgroup (512 threads = 8 wave64) owns one H x W image at a time; the
// image lives in LDS for both 1-D passes, so a whole image-convolution costs
// one HBM read and one HBM write of the image (plus whatever the fused
// prologue / epilogue touches).  Each 1-D pass is a GEMM
//     OUT[r][j] = sum_k IN[r][k] * Kg(k, j),   Kg(k, j) = table[|k - j|]
// run on v_mfma_f32_16x16x4_f32, and its result is written back to LDS
// transposed, so pass 2 is the same routine applied to the transposed
// intermediate (R = K_H . U . K_W with K symmetric, kernels.h:669-699).
//
// The Gaussian factor is never materialised: the B operand is read from the
// 1-D table of exp(-d^2/gamma) kept in LDS.  Each MFMA covers four contiguous
// contraction indices, and with `band` on the 4-wide steps whose factors are
// all exactly 0.0f in FP32 (|k - j| beyond the FP32 underflow of
// exp(-d^2/gamma)) for a unit's 16 output columns are skipped; they would add
// exact zeros.
#pragma once

#include <hip/hip_runtime.h>

namespace wdl {

constexpr int kThreads = 512;              // 8 wave64 per workgroup
constexpr int kWaves = kThreads / 64;
constexpr int kMaxSide = 128;              // largest padded side held in LDS
constexpr int kTabOff = kMaxSide + 16;     // table index of d = 0
constexpr int kTabLen = 2 * kMaxSide + 48; // multiple of 4
constexpr int kTabFloats = kTabLen;        // LDS floats of the FP32 factor table
constexpr int kMaxUnits = 2;               // 16x64 output units per wave per pass

typedef float floatx4 __attribute__((ext_vector_type(4)));

struct ConvGeom {
  int W, H;     // image (x fastest)
  int Wp, Hp;   // padded to a multiple of 16
  int S1;       // LDS row stride of the row-major image (floats)
  int S2;       // LDS row stride of the transposed intermediate (floats)
  int D;        // band half-width: factors with |d| > D are exactly 0
};

// Bytes of dynamic LDS for one workgroup.
__host__ __device__ inline int conv_lds_bytes(const ConvGeom& g) {
  const int img = (g.Hp * g.S1 > g.Wp * g.S2) ? g.Hp * g.S1 : g.Wp * g.S2;
  return (kTabFloats + img) * (int)sizeof(float);
}

// Output units (16 x 64 blocks) per wave when the workgroup has NW waves:
// enough for a 128 x 128 pass.
template <int NW> struct Units {
  static constexpr int n = (kMaxUnits * kWaves + NW - 1) / NW;
};

// The FP32 factor table into LDS: tabs[kTabOff + d] = tab[|d|] (NT threads).
template <int NT = kThreads>
__device__ inline void load_tables(float* tabs, const float* __restrict__ gtab) {
  for (int e = threadIdx.x; e < kTabFloats; e += NT) tabs[e] = gtab[e];
}

// The wave's index in the workgroup as a scalar.  The compiler cannot prove
// threadIdx.x / 64 uniform, and with a per-lane wave index every condition
// derived from it (which output unit, how many row blocks) becomes an
// exec-mask branch around each MFMA, with a wait on all LDS reads before it.
__device__ inline int wave_id() { return __builtin_amdgcn_readfirstlane(threadIdx.x >> 6); }

__device__ inline floatx4 mfma4(float a, float b, floatx4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// One output unit: NRB row blocks of 16 rows, contraction steps ks_lo..ks_hi.
// A step is one MFMA per row block over the four contiguous contraction
// indices 4ks..4ks+3 (lane (c, g) gives A[c][g] = IN[r0 + 16rb + c][4ks + g]
// and B[g][c] = Kg(4ks + g, c0 + c) = tab[abase + 4ks], abase = kTabOff + g -
// c0 - c), so the band is skipped to within 4 indices.  The next step's
// operands are read from LDS while the current step's MFMAs issue; no
// condition inside the loop.
template <int NRB>
__device__ inline void gemm_unit(const float* __restrict__ arow, int sin,
                                 const float* __restrict__ tab, int abase, int ks_lo, int ks_hi,
                                 floatx4 (&acc)[4]) {
  auto load = [&](int ks, float (&a)[NRB], float& b) {
    b = tab[abase + 4 * ks];
#pragma unroll
    for (int rb = 0; rb < NRB; ++rb) a[rb] = arow[rb * 16 * sin + 4 * ks];
  };
  float av[NRB], bv;
  load(ks_lo, av, bv);
  for (int ks = ks_lo; ks <= ks_hi; ++ks) {
    float an[NRB], bn;
    load(min(ks + 1, ks_hi), an, bn);  // the last step re-reads its own operands
#pragma unroll
    for (int rb = 0; rb < NRB; ++rb) acc[rb] = mfma4(av[rb], bv, acc[rb]);
#pragma unroll
    for (int rb = 0; rb < NRB; ++rb) av[rb] = an[rb];
    bv = bn;
  }
}

// One 1-D pass over an LDS matrix IN (nrows x ncontract, row stride sin):
// OUT[r][j] = sum_k IN[r][k] * Kg(k, j + off) for j < nout, Kg(k, j) =
// tab[|k - j|].  Output unit u of this wave covers output columns
// [16*cb, 16*cb+16) and rows [64*rg, 64*rg+64); acc[uu][rb] holds the 16x16
// block of rows 64*rg+16*rb.. in the MFMA 16x16 layout (lane l, reg r: row
// 4*(l/16)+r, col l%16).  Contraction steps whose factors are all beyond D
// (exactly 0 in FP32) are skipped.  NW: waves in the workgroup.
template <int NW = kWaves>
__device__ inline void rows_gemm(const float* __restrict__ in, int sin, int nrows, int ncontract,
                                 int nout, int off, const float* __restrict__ tabs, int D,
                                 floatx4 (&acc)[Units<NW>::n][4]) {
  const int lane = threadIdx.x & 63, wave = wave_id();
  const int c = lane & 15, g = lane >> 4;
  const int ncb = nout >> 4;
  const int nk4 = ncontract >> 2;
  const int nunits = ncb * ((nrows + 63) >> 6);
#pragma unroll
  for (int uu = 0; uu < Units<NW>::n; ++uu) {
#pragma unroll
    for (int rb = 0; rb < 4; ++rb) acc[uu][rb] = floatx4{0.f, 0.f, 0.f, 0.f};
    const int u = wave + uu * NW;
    if (u < nunits) {
      const int cb = u % ncb, rg = u / ncb;
      const int c0 = (cb << 4) + off, r0 = rg << 6;  // c0: contraction index of output column 0
      const int nrb = min(4, (nrows - r0) >> 4);
      // contraction steps of 4 holding a factor that can be nonzero for the
      // unit's 16 columns: indices c0 - D .. c0 + 15 + D
      const int lo = c0 - D;
      const int ks_lo = lo > 0 ? (lo >> 2) : 0;
      const int ks_hi = min(nk4 - 1, (c0 + 15 + D) >> 2);
      const float* arow = in + (r0 + c) * sin + g;
      const int abase = kTabOff + g - c0 - c;
      switch (nrb) {  // wave-uniform
        case 4: gemm_unit<4>(arow, sin, tabs, abase, ks_lo, ks_hi, acc[uu]); break;
        case 3: gemm_unit<3>(arow, sin, tabs, abase, ks_lo, ks_hi, acc[uu]); break;
        case 2: gemm_unit<2>(arow, sin, tabs, abase, ks_lo, ks_hi, acc[uu]); break;
        default: gemm_unit<1>(arow, sin, tabs, abase, ks_lo, ks_hi, acc[uu]); break;
      }
    }
  }
}

// The square pass of the image-resident convolution (nout = ncols, no offset).
template <int NW = kWaves>
__device__ inline void rows_gemm(const float* __restrict__ in, int sin, int nrows, int ncols,
                                 const float* __restrict__ tabs, int D,
                                 floatx4 (&acc)[Units<NW>::n][4]) {
  rows_gemm<NW>(in, sin, nrows, ncols, ncols, 0, tabs, D, acc);
}

// Writes OUT (nrows x ncols output columns, held in acc by rows_gemm) to dst
// transposed: dst[j * sdst + r] = OUT[r][j].
template <int NW = kWaves>
__device__ inline void store_transposed(float* dst, int sdst, int nrows, int ncols,
                                        const floatx4 (&acc)[Units<NW>::n][4]) {
  const int lane = threadIdx.x & 63, wave = wave_id();
  const int c = lane & 15, g = lane >> 4;
  const int ncb = ncols >> 4;
  const int nunits = ncb * ((nrows + 63) >> 6);
#pragma unroll
  for (int uu = 0; uu < Units<NW>::n; ++uu) {
    const int u = wave + uu * NW;
    if (u < nunits) {
      const int cb = u % ncb, rg = u / ncb;
      const int c0 = cb << 4, r0 = rg << 6;
      const int nrb = min(4, (nrows - r0) >> 4);
#pragma unroll
      for (int rb = 0; rb < 4; ++rb)
        if (rb < nrb)
          *reinterpret_cast<floatx4*>(dst + (c0 + c) * sdst + r0 + 16 * rb + 4 * g) = acc[uu][rb];
    }
  }
}

// ---- split-BF16x3 variant -------------------------------------------------------
// Each FP32 operand x is split into bf16 hi = RN(x) and lo = RN(x - hi) (x - hi
// - lo is within 2^-17 |x|), and a product is formed from three bf16 MFMAs,
// hi.hi + hi.lo + lo.hi, accumulated in FP32 (v_mfma_f32_16x16x32_bf16: 16x16
// outputs over 32 contraction indices; lane l holds A[l%16][8(l/16)..+7] and
// B[8(l/16)..+7][l%16]).  The dropped lo.lo term and the split leave ~2^-16
// relative error per product -- between FP32 and BF16, at a third of the
// matrix-core time per FLOP of the FP32 MFMA.  Contraction lengths are
// multiples of 32 (images padded to 32, make_geom).
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
constexpr int kTabBfCopies = 8;  // shifted copies: 8 consecutive bf16 = one 16-byte read
constexpr int kTabBfBytes = 2 * kTabBfCopies * kTabLen * 2;  // hi and lo copies

// copy s of hi (lo) holds the split of tab[x + s] at tb[s*kTabLen + x]
// (tb[(8 + s)*kTabLen + x]).
template <int NT = kThreads>
__device__ inline void load_tables_bf3(__bf16* tb, const float* __restrict__ gtab) {
  for (int e = threadIdx.x; e < kTabBfCopies * kTabLen; e += NT) {
    const int s = e / kTabLen, x = e - s * kTabLen;
    const float v = (x + s < kTabLen) ? gtab[x + s] : 0.f;
    const __bf16 h = (__bf16)v;
    tb[e] = h;
    tb[kTabBfCopies * kTabLen + e] = (__bf16)(v - (float)h);
  }
}

__device__ inline floatx4 mfma_bf16(bf16x8 a, bf16x8 b, floatx4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// rows_gemm with the split-BF16x3 products; same output layout in acc.
template <int NW = kWaves>
__device__ inline void rows_gemm_bf3(const float* __restrict__ in, int sin, int nrows,
                                     int ncontract, int nout, int off,
                                     const __bf16* __restrict__ tb, int D,
                                     floatx4 (&acc)[Units<NW>::n][4]) {
  const int lane = threadIdx.x & 63, wave = wave_id();
  const int c = lane & 15, g = lane >> 4;
  const int ncb = nout >> 4;
  const int nck = ncontract >> 5;
  const int nunits = ncb * ((nrows + 63) >> 6);
#pragma unroll
  for (int uu = 0; uu < Units<NW>::n; ++uu) {
#pragma unroll
    for (int rb = 0; rb < 4; ++rb) acc[uu][rb] = floatx4{0.f, 0.f, 0.f, 0.f};
    const int u = wave + uu * NW;
    if (u < nunits) {
      const int cb = u % ncb, rg = u / ncb;
      const int c0 = (cb << 4) + off, r0 = rg << 6;
      const int nrb = min(4, (nrows - r0) >> 4);
      const int lo = c0 - D;
      const int kc_lo = lo > 0 ? (lo >> 5) : 0;
      const int kc_hi = min(nck - 1, (c0 + 15 + D) >> 5);
      const float* arow = in + (r0 + c) * sin + 8 * g;
      for (int kc = kc_lo; kc <= kc_hi; ++kc) {
        const int kk = kc << 5;
        // B[k][j] = Kg(kk + 8g + t, c0 + c) = tab[kTabOff + kk + 8g + t - c0 - c], t = 0..7
        const int a = kTabOff + kk + 8 * g - c0 - c;
        const int sh = (a & 7) * kTabLen + (a & ~7);
        const bf16x8 bh = *reinterpret_cast<const bf16x8*>(tb + sh);
        const bf16x8 bl = *reinterpret_cast<const bf16x8*>(tb + kTabBfCopies * kTabLen + sh);
#pragma unroll
        for (int rb = 0; rb < 4; ++rb) {
          if (rb >= nrb) continue;
          const float* ap = arow + rb * 16 * sin + kk;
          const floatx4 x0 = *reinterpret_cast<const floatx4*>(ap);
          const floatx4 x1 = *reinterpret_cast<const floatx4*>(ap + 4);
          bf16x8 ah, al;
#pragma unroll
          for (int t = 0; t < 4; ++t) {
            ah[t] = (__bf16)x0[t];
            al[t] = (__bf16)(x0[t] - (float)ah[t]);
            ah[t + 4] = (__bf16)x1[t];
            al[t + 4] = (__bf16)(x1[t] - (float)ah[t + 4]);
          }
          acc[uu][rb] = mfma_bf16(al, bh, acc[uu][rb]);
          acc[uu][rb] = mfma_bf16(ah, bl, acc[uu][rb]);
          acc[uu][rb] = mfma_bf16(ah, bh, acc[uu][rb]);
        }
      }
    }
  }
}

// buf holds U (Hp x Wp, stride S1, zero outside H x W) on entry and
// R = K_H . U . K_W (stride S1; values outside H x W are don't-care) on exit.
// All threads of the workgroup (NW waves) must call it; it synchronises
// internally.
template <int NW = kWaves>
__device__ inline void conv_lds(float* buf, const float* tabs, const ConvGeom& G) {
  floatx4 acc[Units<NW>::n][4];
  __syncthreads();
  rows_gemm<NW>(buf, G.S1, G.Hp, G.Wp, tabs, G.D, acc);   // T = U . K_W
  __syncthreads();
  store_transposed<NW>(buf, G.S2, G.Hp, G.Wp, acc);        // buf = T^T (Wp x Hp)
  __syncthreads();
  rows_gemm<NW>(buf, G.S2, G.Wp, G.Hp, tabs, G.D, acc);   // R^T = T^T . K_H
  __syncthreads();
  store_transposed<NW>(buf, G.S1, G.Wp, G.Hp, acc);        // buf = R (Hp x Wp)
  __syncthreads();
}

// conv_lds with the split-BF16x3 passes (Wp, Hp multiples of 32).
template <int NW = kWaves>
__device__ inline void conv_lds_bf3(float* buf, const __bf16* tb, const ConvGeom& G) {
  floatx4 acc[Units<NW>::n][4];
  __syncthreads();
  rows_gemm_bf3<NW>(buf, G.S1, G.Hp, G.Wp, G.Wp, 0, tb, G.D, acc);
  __syncthreads();
  store_transposed<NW>(buf, G.S2, G.Hp, G.Wp, acc);
  __syncthreads();
  rows_gemm_bf3<NW>(buf, G.S2, G.Wp, G.Hp, G.Hp, 0, tb, G.D, acc);
  __syncthreads();
  store_transposed<NW>(buf, G.S1, G.Wp, G.Hp, acc);
  __syncthreads();
}

// The table region and convolution of one precision variant: BF3 = false is
// FP32 MFMA (4 shifted FP32 copies), true the split-BF16x3 form.
template <bool BF3> struct ConvPrec;
template <> struct ConvPrec<false> {
  static constexpr int kTabBytes = kTabFloats * (int)sizeof(float);
  template <int NT = kThreads>
  __device__ static void load(void* t, const float* g) { load_tables<NT>((float*)t, g); }
  template <int NW = kWaves>
  __device__ static void conv(float* buf, const void* t, const ConvGeom& G) {
    conv_lds<NW>(buf, (const float*)t, G);
  }
};
template <> struct ConvPrec<true> {
  static constexpr int kTabBytes = kTabBfBytes;
  template <int NT = kThreads>
  __device__ static void load(void* t, const float* g) { load_tables_bf3<NT>((__bf16*)t, g); }
  template <int NW = kWaves>
  __device__ static void conv(float* buf, const void* t, const ConvGeom& G) {
    conv_lds_bf3<NW>(buf, (const __bf16*)t, G);
  }
};

// Visits every padded position of the LDS image; one wave per row, so global
// accesses of a wave are 64 consecutive floats.  f(lds_index, pixel, valid).
template <class F>
__device__ inline void for_each_px(const ConvGeom& G, F&& f) {
  const int lane = threadIdx.x & 63, wave = wave_id();
  for (int row = wave; row < G.Hp; row += kWaves)
    for (int col = lane; col < G.Wp; col += 64)
      f(row * G.S1 + col, row * G.W + col, row < G.H && col < G.W);
}

// ---- vectorised elementwise passes -------------------------------------------
// Elementwise steps touch VW consecutive pixels per call (VW = 4 when W % 4 == 0,
// so a group never straddles a row, else 1).  Each thread first issues the
// global loads of kUnroll groups, then applies them: the loads of one thread
// are in flight together instead of one dependent chain at a time.
template <int VW> struct Vec;
template <> struct Vec<1> { typedef float T; };
template <> struct Vec<4> { typedef floatx4 T; };

template <int VW>
__device__ inline typename Vec<VW>::T vld(const float* p) {
  return *reinterpret_cast<const typename Vec<VW>::T*>(p);
}
template <int VW>
__device__ inline void vst(float* p, typename Vec<VW>::T v) {
  *reinterpret_cast<typename Vec<VW>::T*>(p) = v;
}
template <int VW>
__device__ inline typename Vec<VW>::T vsplat(float x) {
  return typename Vec<VW>::T(x);
}

// Visits load(j) for j = 0..K-1 in ascending j as f(j, value), with the loads
// of B atoms in flight together: K is a runtime value, and a plain loop over it
// waits for each load before issuing the next.
constexpr int kAtomBatch = 8;
template <int B = kAtomBatch, class Load, class F>
__device__ inline void for_atoms_l(int K, Load&& load, F&& f) {
  for (int j0 = 0; j0 < K; j0 += B) {
    decltype(load(0)) v[B];
#pragma unroll
    for (int t = 0; t < B; ++t)
      if (j0 + t < K) v[t] = load(j0 + t);
#pragma unroll
    for (int t = 0; t < B; ++t)
      if (j0 + t < K) f(j0 + t, v[t]);
  }
}
// The same over x[j * stride] (FP32 groups of VW pixels).
template <int VW, class F>
__device__ inline void for_atoms(const float* __restrict__ x, long long stride, int K, F&& f) {
  for_atoms_l(K, [&](int j) { return vld<VW>(x + (long long)j * stride); }, f);
}

template <class F> __device__ inline float vmap(float x, F f) { return f(x); }
template <class F> __device__ inline floatx4 vmap(floatx4 x, F f) {
  return floatx4{f(x.x), f(x.y), f(x.z), f(x.w)};
}
__device__ inline double vdot_d(float a, float b) { return (double)a * (double)b; }
__device__ inline double vdot_d(floatx4 a, floatx4 b) {
  return (double)a.x * b.x + (double)a.y * b.y + (double)a.z * b.z + (double)a.w * b.w;
}

constexpr int kUnroll = 4;
// px_pass's default: measured on the resident 128x128 stages, 2 groups per thread
// in flight beat 1, 4 and 8 (more registers in flight slowed the backward stage).
constexpr int kUnrollPx = 2;

// load(gi) -> any value type; apply(li, gi, valid, value).  li / gi are the
// LDS and pixel indices of the group's first pixel.  U groups per thread have
// their loads in flight together (passes that load few arrays take more).
// NT: threads in the workgroup.
template <int VW, int U = kUnrollPx, int NT = kThreads, class Load, class Apply>
__device__ inline void px_pass(const ConvGeom& G, Load&& load, Apply&& apply) {
  const int q = G.Wp / VW;
  const int total = G.Hp * q;
  typedef decltype(load(0)) V;
  // Group e = threadIdx.x + k kThreads sits at row e / q, column VW (e % q);
  // a thread's successive groups advance by (drow, dcol) rows and groups, plus
  // one row when the column wraps -- no per-group integer division.
  const int drow = NT / q, dcol = NT - drow * q;
  int row = threadIdx.x / q, col = threadIdx.x - row * q;
  int lrow = row * G.S1 + col * VW, grow = row * G.W + col * VW;
  for (int base = threadIdx.x; base < total; base += NT * U) {
    V v[U];
    int li[U], gi[U];
    bool ok[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      li[u] = lrow;
      gi[u] = grow;
      ok[u] = row < G.H && col * VW < G.W;
      if (ok[u]) v[u] = load(gi[u]);
      row += drow;
      col += dcol;
      lrow += drow * G.S1 + dcol * VW;
      grow += drow * G.W + dcol * VW;
      if (col >= q) {
        col -= q;
        ++row;
        lrow += G.S1 - G.Wp;
        grow += G.W - G.Wp;
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (base + u * NT < total) apply(li[u], gi[u], ok[u], v[u]);
  }
}

// Wave scans with DPP lane moves (VALU; a __shfl is one LDS-crossbar round
// trip per step).  Row shifts by 1, 2, 4, 8 within each row of 16 lanes, then
// the broadcast of lane 15 into the next row (half scan: lanes 31 and 63 hold
// their half's result), then of lane 31 into the upper half (lane i holds the
// result over lanes 0..i, lane 63 the wave's).  A lane whose source lies
// outside its row, or whose row a broadcast leaves out, keeps its own value,
// which max and min absorb.  Every lane of the wave must be active.
template <int CTRL, int ROWS>
__device__ inline int dpp_move(int v) {
  return __builtin_amdgcn_update_dpp(v, v, CTRL, ROWS, 0xf, false);
}
template <int CTRL, int ROWS>
__device__ inline float dpp_move(float v) {
  return __int_as_float(dpp_move<CTRL, ROWS>(__float_as_int(v)));
}
template <class T, class Op>
__device__ inline T half_scan(T x, Op op) {
  x = op(x, dpp_move<0x111, 0xf>(x));  // row_shr:1
  x = op(x, dpp_move<0x112, 0xf>(x));  // row_shr:2
  x = op(x, dpp_move<0x114, 0xf>(x));  // row_shr:4
  x = op(x, dpp_move<0x118, 0xf>(x));  // row_shr:8
  return op(x, dpp_move<0x142, 0xa>(x));  // row_bcast:15
}
template <class T, class Op>
__device__ inline T wave_scan(T x, Op op) {
  const T h = half_scan(x, op);
  return op(h, dpp_move<0x143, 0xc>(h));  // row_bcast:31 after the halves
}
struct MaxOp {
  __device__ int operator()(int a, int b) const { return max(a, b); }
  __device__ float operator()(float a, float b) const { return fmaxf(a, b); }
};
struct MinOp {
  __device__ float operator()(float a, float b) const { return fminf(a, b); }
};
// The wave's maximum / minimum in every lane.
__device__ inline float wave_max(float x) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(wave_scan(x, MaxOp())), 63));
}
__device__ inline float wave_min(float x) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(wave_scan(x, MinOp())), 63));
}

// safe_log (loss.h:3-5) in FP32: log(max(1e-250, x)); 1e-250 is below the
// FP32 range, so every x <= 0 (and NaN, as std::max returns its first
// argument) maps to log(1e-250).
// log and exp of the elementwise steps on the hardware log2 / exp2, without the
// library's range reductions.  flog adds the exact binary exponent to the
// mantissa's log2 (in [-1, 0): within ~1e-7 absolute of log x, plus the final
// rounding).  fexp multiplies the hardware 2^h, h = log2(e) x rounded, by
// 1 + r ln 2, r the rounding's remainder (exact by FMA) plus the constant's own
// error: within about 2 ulp of exp x (below 2^-126 the hardware may flush).
__device__ inline float flog(float x) {
  int e;
  const float m = frexpf(x, &e);
  return fmaf((float)e, 0.693147182f, __builtin_amdgcn_logf(m) * 0.693147182f);
}
__device__ inline float fexp(float x) {
  const float h = x * 1.44269502f;
  const float r = fmaf(x, 1.44269502f, -h) + x * 1.92596299e-8f;
  const float e = __builtin_amdgcn_exp2f(h);
  // at the ends of the range (2^h beyond 2^+-126, inf, 0, NaN) 2^h alone
  return fabsf(h) < 126.f ? fmaf(e, r * 0.693147182f, e) : e;
}
__device__ inline float safe_logf(float x) { return x > 0.f ? flog(x) : -575.64627324851142f; }
__device__ inline floatx4 safe_logf(floatx4 x) {
  return floatx4{safe_logf(x.x), safe_logf(x.y), safe_logf(x.z), safe_logf(x.w)};
}
// x / y of the elementwise steps: the hardware reciprocal (1 ulp) and one
// Newton-Raphson correction of the quotient, q1 = q0 + r (x - y q0) with
// q0 = x r (Markstein, IBM J. Res. Develop. 34(1), 1990): within 1 ulp, and
// mostly the rounded quotient, when y and x / y are normal FP32 numbers --
// 4 instructions against ~10 for the IEEE expansion.  Out of that range
// (y = 0, +-inf, subnormal) the result may be NaN where IEEE gives +-inf or a
// huge finite value; such scalings have left the FP32 range anyway, and the
// solver counts them as range events.
__device__ inline float fdiv(float x, float y) {
  const float r = __builtin_amdgcn_rcpf(y);
  const float q = x * r;
  return fmaf(fmaf(-y, q, x), r, q);
}
__device__ inline floatx4 fdiv(floatx4 x, floatx4 y) {
  return floatx4{fdiv(x.x, y.x), fdiv(x.y, y.y), fdiv(x.z, y.z), fdiv(x.w, y.w)};
}
__device__ inline floatx4 fdiv(float x, floatx4 y) {
  return floatx4{fdiv(x, y.x), fdiv(x, y.y), fdiv(x, y.z), fdiv(x, y.w)};
}
// Natural log from the hardware log2 (v_log_f32, about 1 ulp) for values that
// only pick a shift or a range key (the shifted passes' line maxima: the result
// does not depend on the shift's exact value).  The instruction flushes
// denormal inputs, so values below FLT_MIN are scaled by 2^64 first.
__device__ inline float fast_logf(float x) {
  const bool small = x < 1.17549435e-38f;
  const float y = small ? x * 18446744073709551616.f : x;
  return (__builtin_amdgcn_logf(y) - (small ? 64.f : 0.f)) * 0.693147180559945309f;
}
__device__ inline float vexpf(float x) { return fexp(x); }
__device__ inline floatx4 vexpf(floatx4 x) { return floatx4{fexp(x.x), fexp(x.y), fexp(x.z), fexp(x.w)}; }
__device__ inline float vlogf(float x) { return flog(x); }
__device__ inline floatx4 vlogf(floatx4 x) { return floatx4{flog(x.x), flog(x.y), flog(x.z), flog(x.w)}; }

}  // namespace wdl
