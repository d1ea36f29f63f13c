// Tiled Sinkhorn stages for images with a side above 128, which no longer fit
// one workgroup's LDS.  Same algorithm, precision and device layout as the
// LDS-resident stages in wdl_solver.hip (compute_gradient_both,
// inverseWasserstein.h:544-674; compute_gradient_both_log_signArray,
// :678-814); what changes is how a stage is cut into launches:
//   forward s   (standard)  Qh[s-2] = exp(sum_j lam_j safe_log phi_j^{s-1})   k_tq
//                           c^{s-1} = K b^{s-1},  b^{s-1} = q^{s-1} / phi^{s-1}  k_tfwd1
//                           phi^s   = K (p / c^{s-1})                           k_tfwd2
//   backward s              first convolution of both chains, their fused
//                           epilogues, and the input of the second (X0 / X1)   k_tbwd1
//                           second convolution and the next v / r               k_tbwd2
// (the log domain computes log b in k_tfwd1 from the FP64 log phi directly).
// A workgroup owns one 64 x 64 output tile of one (histogram, atom) image: it
// stages the tile's haloed region in LDS and convolves it (wdl_tile.h).
// Elementwise steps are evaluated on the haloed region for the convolution
// input and on the inner tile for outputs, so every stored value is written by
// exactly one workgroup.  The weight-gradient dot of a (histogram, atom) is
// accumulated per tile and summed over the tiles at the end (k_tgl); its
// centring constant (the image mean of log phi, see exact_gsum) comes from
// per-tile sums the forward writes into Mt.
#include <atomic>
#include <algorithm>
#include <cmath>

#include "wdl_solve_common.h"

namespace wdl {

// Workgroup -> (histogram, atom, tile).  A unit is a (histogram, tile) pair;
// its K workgroups (one per atom) take consecutive slots of one XCD, and
// every XCD takes a contiguous range of units, so the K atoms of a tile, and
// neighbouring tiles, share their reads in that XCD's L2.
__device__ inline bool map_tile_unit(const SolveArgs& a, int slot, int xcd, int& p, int& i,
                                     int& tile) {
  const int nt = a.T.nx * a.T.ny;
  const int nunits = a.np * nt;
  const int per = (nunits + 7) >> 3;
  const int uo = slot / a.K;
  if (uo >= per) return false;
  const int u = xcd * per + uo;
  if (u >= nunits) return false;
  i = slot - uo * a.K;
  const int pl = u / nt;
  p = a.p0 + pl;
  tile = u - pl * nt;
  return true;
}

__device__ inline bool map_tile(const SolveArgs& a, int& p, int& i, int& tile) {
  return map_tile_unit(a, (int)(blockIdx.x >> 3), (int)(blockIdx.x & 7), p, i, tile);
}

// Backward: the two chains of a (histogram, atom, tile) on adjacent slots.
__device__ inline bool map_tile_bwd(const SolveArgs& a, int& p, int& i, int& tile, int& chain) {
  if (a.weightsOnly) {
    chain = 1;
    return map_tile(a, p, i, tile);
  }
  const int slot = (int)(blockIdx.x >> 3);
  chain = slot & 1;
  return map_tile_unit(a, slot >> 1, (int)(blockIdx.x & 7), p, i, tile);
}

// Slot of tile `tile` of (histogram p, atom i) at iteration s in the per-tile
// sums of log phi, and the image mean assembled from them.
__device__ inline long long mt_index(const SolveArgs& a, int s, int p, int i, int tile) {
  return (((long long)(s - 1) * a.P + p) * a.K + i) * (a.T.nx * a.T.ny) + tile;
}
__device__ inline double mean_log_phi(const SolveArgs& a, int s, int p, int i) {
  const int nt = a.T.nx * a.T.ny;
  const double* m = a.Mt + mt_index(a, s, p, i, 0);
  double t = 0;
  for (int k = 0; k < nt; ++k) t += m[k];
  return t / (double)a.N;
}

static int tile_grid(const SolveArgs& a) {
  const long long nunits = (long long)a.np * a.T.nx * a.T.ny;
  return (int)(8LL * a.K * ((nunits + 7) / 8));
}

struct TileSmem {
  float* tabs;
  float* buf;
  LogLds L;
};

// Log domain: the row keys (R region rows), the column keys (64 inner
// columns) and the finite minima of either, right after the image: a
// 96-row region then takes 40 KB of LDS in all (four workgroups per CU), an
// 80-row one 29 KB.
__host__ __device__ inline int tile_keys(const TileGeom& T) {
  return T.R + kTile + (T.R > kTile ? T.R : kTile);
}

__device__ inline TileSmem tile_smem(float* smem, const TileGeom& T) {
  TileSmem m;
  m.tabs = smem;
  m.buf = smem + kTabFloats;
  int* keys = reinterpret_cast<int*>(m.buf + T.R * T.S);
  m.L.mrow = keys;
  m.L.mcol = keys + T.R;
  m.L.rmin = keys + T.R + kTile;
  return m;
}

// ---- standard domain -----------------------------------------------------------

// q^{s-1} of every histogram that runs iteration s (:577-585); k_final writes q^{L_p}.
template <int VW>
__global__ __launch_bounds__(256) void k_tq(SolveArgs a, int s) {
  typedef typename Vec<VW>::T V;
  const int p = a.p0 + (int)blockIdx.y;
  if (!fwd_runs(a, p, s)) return;
  const long long N = a.N;
  const int K = a.K;
  const float* F = a.Fh + hslot(a, s - 2) * a.stage + (long long)p * K * N;
  float* Q = a.Qh + hslot(a, s - 2) * a.qstage + (long long)p * N;
  const float* lam = a.lam + (long long)p * K;
  for (long long x = ((long long)blockIdx.x * blockDim.x + threadIdx.x) * VW; x < N;
       x += (long long)gridDim.x * blockDim.x * VW) {
    V lq = vsplat<VW>(0.f);
    for_atoms<VW>(F + x, N, K, [&](int j, V f) { lq += lam[j] * safe_logf(f); });
    vst<VW>(Q + x, vexpf(lq));
  }
}

// c^{s-1} = K b^{s-1} (:568), b^0 = 1 or the warm-start store, else q / phi (:587-592).
template <int VW>
__global__ __launch_bounds__(kThreads, 4) void k_tfwd1(SolveArgs a, int s) {
  typedef typename Vec<VW>::T V;
  int p, i, tile;
  if (!map_tile(a, p, i, tile) || !fwd_runs(a, p, s)) return;
  extern __shared__ __align__(16) float smem[];
  const TileGeom T = a.T;
  const TileSmem sm = tile_smem(smem, T);
  load_tables(sm.tabs, a.tab);
  const int tx = tile % T.nx, ty = tile / T.nx;
  const long long N = a.N;
  const long long img = ((long long)p * a.K + i) * N;
  if (s == 1) {
    const float* bi = a.bInit ? a.bInit + img : nullptr;
    tile_region_pass<VW>(
        T, tx, ty, [&](int gi) { return bi ? vld<VW>(bi + gi) : vsplat<VW>(1.f); },
        [&](int li, int, bool ok, V b) { vst<VW>(sm.buf + li, ok ? b : vsplat<VW>(0.f)); });
  } else {
    const float* Qp = a.Qh + hslot(a, s - 2) * a.qstage + (long long)p * N;  // q^{s-1}
    const float* Fp = a.Fh + hslot(a, s - 2) * a.stage + img;                // phi^{s-1}
    struct Two { V q, f; };
    tile_region_pass<VW>(
        T, tx, ty,
        [&](int gi) {
          Two r;
          r.q = vld<VW>(Qp + gi);
          r.f = vld<VW>(Fp + gi);
          return r;
        },
        [&](int li, int, bool ok, const Two& r) {
          vst<VW>(sm.buf + li, ok ? fdiv(r.q, r.f) : vsplat<VW>(0.f));
        });
  }
  tile_conv(sm.buf, sm.tabs, T);
  float* Cd = a.Ch + hslot(a, s - 1) * a.stage + img;
  tile_inner_pass<VW>(
      T, tx, ty, [&](int) { return 0; },
      [&](int li, int gi, bool ok, int) {
        if (ok) vst<VW>(Cd + gi, vld<VW>(sm.buf + li));
      });
}

// phi^s = K a, a = p / c^{s-1} (:569-575).
template <int VW>
__global__ __launch_bounds__(kThreads, 4) void k_tfwd2(SolveArgs a, int s) {
  typedef typename Vec<VW>::T V;
  int p, i, tile;
  if (!map_tile(a, p, i, tile) || !fwd_runs(a, p, s)) return;
  extern __shared__ __align__(16) float smem[];
  const TileGeom T = a.T;
  const TileSmem sm = tile_smem(smem, T);
  load_tables(sm.tabs, a.tab);
  const int tx = tile % T.nx, ty = tile / T.nx;
  const long long N = a.N;
  const long long img = ((long long)p * a.K + i) * N;
  const float* Ai = a.A + (long long)i * N;
  const float* Cs = a.Ch + hslot(a, s - 1) * a.stage + img;
  struct Two { V a, c; };
  tile_region_pass<VW>(
      T, tx, ty,
      [&](int gi) {
        Two r;
        r.a = vld<VW>(Ai + gi);
        r.c = vld<VW>(Cs + gi);
        return r;
      },
      [&](int li, int, bool ok, const Two& r) {
        vst<VW>(sm.buf + li, ok ? fdiv(r.a, r.c) : vsplat<VW>(0.f));
      });
  tile_conv(sm.buf, sm.tabs, T);
  float* Fd = a.Fh + hslot(a, s - 1) * a.stage + img;
  double lsum = 0;  // sum of log phi over the tile, for the centred weight-gradient dot
  tile_inner_pass<VW>(
      T, tx, ty, [&](int) { return 0; },
      [&](int li, int gi, bool ok, int) {
        if (!ok) return;
        const V f = vld<VW>(sm.buf + li);
        vst<VW>(Fd + gi, f);
        lsum += vsum_d(vlogf(f));
      });
  const double t = block_sum(lsum);
  if (threadIdx.x == 0) a.Mt[mt_index(a, s, p, i, tile)] = t;
}

// Backward step s, first convolution.  chain 0 (:604-616): c = K((lam_i n - v_i) b^s),
// resultDic += c / c^{s-1}, X0 = -p c / (c^{s-1})^2.  chain 1 (:640-660): the dot
// <log phi^s, g>, c = K((lam_i g - r_i) / phi^s), X1 = c (-p / (c^{s-1})^2).
template <int VW>
__global__ __launch_bounds__(kThreads, 4) void k_tbwd1(SolveArgs a, int t) {
  typedef typename Vec<VW>::T V;
  int p, i, tile, chain;
  if (!map_tile_bwd(a, p, i, tile, chain)) return;
  const int s = bwd_step(a, p, t);
  if (s < 1) return;
  extern __shared__ __align__(16) float smem[];
  const TileGeom T = a.T;
  const TileSmem sm = tile_smem(smem, T);
  load_tables(sm.tabs, a.tab);
  const int tx = tile % T.nx, ty = tile / T.nx;
  const long long N = a.N;
  const int K = a.K;
  const long long hbase = (long long)p * K * N;
  const long long img = hbase + (long long)i * N;
  const float lpi = a.lam[(long long)p * K + i];
  const float* Ai = a.A + (long long)i * N;
  const float* Cs = a.Ch + hslot(a, s - 1) * a.stage + img;  // c^{s-1}
  const float* Fs = a.Fh + hslot(a, s - 1) * a.stage + img;  // phi^s
  const float* G0 = a.G0 + (long long)p * N;
  const bool last = (s == steps_of(a, p));
  struct Three { V x, y, z; };
  struct Two { V x, y; };
  if (chain == 0) {
    const float* Vp = (s & 1) ? a.V0 : a.V1;
    const float* Qs = a.Qh + hslot(a, s - 1) * a.qstage + (long long)p * N;  // q^s
    tile_region_pass<VW>(
        T, tx, ty,
        [&](int gi) {
          Three r;  // n, v_i, b^s = q^s / phi^s
          r.z = fdiv(vld<VW>(Qs + gi), vld<VW>(Fs + gi));
          if (last) {
            r.x = vld<VW>(G0 + gi);
            r.y = vsplat<VW>(0.f);
          } else {
            r.x = vsplat<VW>(0.f);
            r.y = vsplat<VW>(0.f);
            for_atoms<VW>(Vp + hbase + gi, N, K, [&](int j, V v) {
              r.x += v;
              if (j == i) r.y = v;
            });
          }
          return r;
        },
        [&](int li, int, bool ok, const Three& r) {
          vst<VW>(sm.buf + li, ok ? (lpi * r.x - r.y) * r.z : vsplat<VW>(0.f));
        });
    tile_conv(sm.buf, sm.tabs, T);
    float* GDp = a.GD + img;
    float* X = a.X0 + img;
    tile_inner_pass<VW>(
        T, tx, ty,
        [&](int gi) {
          Three r;  // conv_b, resultDic so far, atom
          r.x = vld<VW>(Cs + gi);
          r.y = last ? vsplat<VW>(0.f) : vld<VW>(GDp + gi);
          r.z = vld<VW>(Ai + gi);
          return r;
        },
        [&](int li, int gi, bool ok, const Three& r) {
          if (!ok) return;
          const V c = vld<VW>(sm.buf + li);
          const V cq = fdiv(c, r.x);
          vst<VW>(GDp + gi, r.y + cq);
          if (s > 1) vst<VW>(X + gi, fdiv(-r.z * cq, r.x));
        });
  } else {
    const float* Rp = (s & 1) ? a.R0 : a.R1;
    const float* Qp = a.Q + (long long)p * N;
    // centred dot <log phi - m, g> (m = 0 at the last step); tile 0 adds m E_s
    const double m = last ? 0.0 : mean_log_phi(a, s, p, i);
    double dot = 0;
    auto wload = [&](int gi) {
      Three r;  // g, r_i, phi^s
      r.z = vld<VW>(Fs + gi);
      if (last) {
        r.x = vld<VW>(G0 + gi) * vld<VW>(Qp + gi);
        r.y = vsplat<VW>(0.f);
      } else {
        r.x = vsplat<VW>(0.f);
        r.y = vsplat<VW>(0.f);
        for_atoms<VW>(Rp + hbase + gi, N, K, [&](int j, V v) {
          r.x += v;
          if (j == i) r.y = v;
        });
      }
      return r;
    };
    auto acc_dot = [&](const Three& r) {
      dot += vdot_d(vlogf(r.z), r.x) - m * vsum_d(r.x);  // plain log (:649)
    };
    if (s == 1) {
      tile_inner_pass<VW>(T, tx, ty, wload, [&](int, int, bool ok, const Three& r) {
        if (ok) acc_dot(r);
      });
    } else {
      tile_region_pass<VW>(T, tx, ty, wload, [&](int li, int, bool ok, const Three& r) {
        V u = vsplat<VW>(0.f);
        if (ok) {
          if (tile_inner(T, li)) acc_dot(r);
          u = fdiv(lpi * r.x - r.y, r.z);
        }
        vst<VW>(sm.buf + li, u);
      });
    }
    double tot = block_sum(dot);
    if (threadIdx.x == 0) {
      if (!last && tile == 0) tot += m * exact_gsum(a, p, s);
      double* gl = a.GLt + ((long long)p * K + i) * (T.nx * T.ny) + tile;
      *gl = last ? tot : *gl + tot;
    }
    if (s == 1) return;
    tile_conv(sm.buf, sm.tabs, T);
    float* X = a.X1 + img;
    tile_inner_pass<VW>(
        T, tx, ty,
        [&](int gi) {
          Two r;
          r.x = vld<VW>(Cs + gi);
          r.y = vld<VW>(Ai + gi);
          return r;
        },
        [&](int li, int gi, bool ok, const Two& r) {
          if (ok) vst<VW>(X + gi, vld<VW>(sm.buf + li) * fdiv(fdiv(-r.y, r.x), r.x));
        });
  }
}

// Backward step s >= 2, second convolution: v^{s-1}_i = K X0 / phi^{s-1} (:617-627),
// r^{s-1}_i = b^{s-1} K X1 (:661-670).
template <int VW>
__global__ __launch_bounds__(kThreads, 4) void k_tbwd2(SolveArgs a, int t) {
  typedef typename Vec<VW>::T V;
  int p, i, tile, chain;
  if (!map_tile_bwd(a, p, i, tile, chain)) return;
  const int s = bwd_step(a, p, t);
  if (s < 2) return;
  extern __shared__ __align__(16) float smem[];
  const TileGeom T = a.T;
  const TileSmem sm = tile_smem(smem, T);
  load_tables(sm.tabs, a.tab);
  const int tx = tile % T.nx, ty = tile / T.nx;
  const long long N = a.N;
  const long long img = ((long long)p * a.K + i) * N;
  const float* X = (chain == 0 ? a.X0 : a.X1) + img;
  tile_region_pass<VW>(
      T, tx, ty, [&](int gi) { return vld<VW>(X + gi); },
      [&](int li, int, bool ok, V x) { vst<VW>(sm.buf + li, ok ? x : vsplat<VW>(0.f)); });
  tile_conv(sm.buf, sm.tabs, T);
  const float* Fm = a.Fh + hslot(a, s - 2) * a.stage + img;  // phi^{s-1}
  if (chain == 0) {
    float* Vc = ((s & 1) ? a.V1 : a.V0) + img;
    tile_inner_pass<VW>(
        T, tx, ty, [&](int gi) { return vld<VW>(Fm + gi); },
        [&](int li, int gi, bool ok, V f) {
          if (ok) vst<VW>(Vc + gi, fdiv(vld<VW>(sm.buf + li), f));
        });
  } else {
    float* Rc = ((s & 1) ? a.R1 : a.R0) + img;
    const float* Qm = a.Qh + hslot(a, s - 2) * a.qstage + (long long)p * N;  // q^{s-1}
    tile_inner_pass<VW>(
        T, tx, ty, [&](int gi) { return fdiv(vld<VW>(Qm + gi), vld<VW>(Fm + gi)); },  // b^{s-1}
        [&](int li, int gi, bool ok, V b) {
          if (ok) vst<VW>(Rc + gi, vld<VW>(sm.buf + li) * b);
        });
  }
}

// resultLa[p][i] = sum over tiles of the per-tile dots.
__global__ void k_tgl(SolveArgs a) {
  const int nt = a.T.nx * a.T.ny;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= a.np * a.K) return;
  const long long pk = (long long)a.p0 * a.K + e;
  double acc = 0;
  for (int t = 0; t < nt; ++t) acc += a.GLt[pk * nt + t];
  a.GL[pk] = acc;
}

// ---- log domain ----------------------------------------------------------------
// Bh/Ch/Fh hold FP64 log b, log conv_b, log phi; X0/X1 the FP64 log
// magnitude of the second backward convolution's input and XS0/XS1 its sign
// (1 = non-negative), one byte per value.

template <int VW> struct SignIO;
template <> struct SignIO<1> {
  __device__ static float load(const unsigned char* p) { return *p ? 1.f : -1.f; }
  __device__ static void store(unsigned char* p, float s) { *p = s > 0.f ? 1 : 0; }
};
template <> struct SignIO<4> {
  __device__ static floatx4 load(const unsigned char* p) {
    const unsigned w = *reinterpret_cast<const unsigned*>(p);
    return floatx4{(w & 0xFFu) ? 1.f : -1.f, (w & 0xFF00u) ? 1.f : -1.f,
                   (w & 0xFF0000u) ? 1.f : -1.f, (w & 0xFF000000u) ? 1.f : -1.f};
  }
  __device__ static void store(unsigned char* p, floatx4 s) {
    *reinterpret_cast<unsigned*>(p) = (s.x > 0.f ? 1u : 0u) | (s.y > 0.f ? 1u << 8 : 0u) |
                                      (s.z > 0.f ? 1u << 16 : 0u) | (s.w > 0.f ? 1u << 24 : 0u);
  }
};

__device__ inline double tdot(double a, float b) { return a * (double)b; }
__device__ inline double tdot(doublex4 a, floatx4 b) {
  return a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w;
}

// Log domain, forward step s >= 2: log b^{s-1} = lq - log phi_i^{s-1} of every
// atom, lq = sum_j lam_j log phi_j^{s-1} (:711-724), into the history Bh[s-2].
// One thread per (histogram, pixel) forms lq once for the K images; the tiles
// of k_tfwd1_log then load one image over their region instead of K.
__global__ __launch_bounds__(256) void k_tlogb(SolveArgs a, int s) {
  const long long N = a.N, tot = (long long)a.np * N;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < tot;
       e += (long long)gridDim.x * blockDim.x) {
    const int p = a.p0 + (int)(e / N);
    const long long x = e % N;
    if (!fwd_runs(a, p, s)) continue;
    const long long hb = (long long)p * a.K * N;
    const double* Fp = (const double*)a.Fh + hslot(a, s - 2) * a.stage + hb + x;
    double* Bp = (double*)a.Bh + hslot(a, s - 2) * a.stage + hb + x;
    const float* lam = a.lam + (long long)p * a.K;
    double lq = 0;
    for_atoms_l<4>(
        a.K, [&](int j) { return Fp[j * N]; },
        [&](int j, double f) { lq += (double)lam[j] * f; });
    for (int i = 0; i < a.K; ++i) Bp[i * N] = lq - Fp[i * N];
  }
}

// Log domain, backward launch t: the cross-atom sums of step s (not the last)
// that every atom's workgroup of k_tbwd1_log needs, n = -sum_j v_j (:764) and
// g = -sum_j r_j (:808), once per (histogram, pixel), summed in atom order as
// the tiles did.
__global__ __launch_bounds__(256) void k_tjoin(SolveArgs a, int t) {
  const long long N = a.N, tot = (long long)a.np * N;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < tot;
       e += (long long)gridDim.x * blockDim.x) {
    const int p = a.p0 + (int)(e / N);
    const long long x = e % N;
    const int s = bwd_step(a, p, t);
    if (s < 1 || s == steps_of(a, p)) continue;
    const long long hb = (long long)p * a.K * N;
    if (!a.weightsOnly) {
      const float* Vp = ((s & 1) ? a.V0 : a.V1) + hb + x;
      float n = 0.f;
      for_atoms_l(a.K, [&](int j) { return Vp[j * N]; }, [&](int, float v) { n -= v; });
      a.nj[(long long)p * N + x] = n;
    }
    const float* Rp = ((s & 1) ? a.R0 : a.R1) + hb + x;
    float g = 0.f;
    for_atoms_l(a.K, [&](int j) { return Rp[j * N]; }, [&](int, float v) { g -= v; });
    a.gj[(long long)p * N + x] = g;
  }
}

// log c^{s-1} = K log b^{s-1}, log b^{s-1} = sum_j lam_j log phi_j^{s-1} - log phi_i^{s-1}
// (:711-724; b^0: 0 or the warm-start store), also stored as the history Bh.
template <int VW>
__global__ __launch_bounds__(kThreads, 6) void k_tfwd1_log(SolveArgs a, int s) {
  typedef typename Vec<VW>::T V;
  typedef typename VecD<VW>::T VD;
  int p, i, tile;
  if (!map_tile(a, p, i, tile) || !fwd_runs(a, p, s)) return;
  extern __shared__ __align__(16) float smem[];
  const TileGeom T = a.T;
  const TileSmem sm = tile_smem(smem, T);
  load_tables(sm.tabs, a.tab);
  const int tx = tile % T.nx, ty = tile / T.nx;
  const long long N = a.N;
  const int K = a.K;
  const long long hbase = (long long)p * K * N;
  const long long img = hbase + (long long)i * N;
  if (s == 1) {
    const double* bi = a.lbInit ? a.lbInit + img : nullptr;  // warm start (log b, FP64)
    tile_log_conv<VW>(
        sm.buf, sm.tabs, sm.L, T, tx, ty, a.events,
        [&](int gi) { return bi ? vldd<VW>(bi + gi) : vd(vsplat<VW>(0.f)); },
        [&](int, int, VD b, bool) { return SLog<VW>{b, vsplat<VW>(1.f)}; });
  } else {
    // log b^{s-1}, written to the history by k_tlogb
    const double* Bd = (const double*)a.Bh + hslot(a, s - 2) * a.stage + img;
    tile_log_conv<VW>(
        sm.buf, sm.tabs, sm.L, T, tx, ty, a.events, [&](int gi) { return vldd<VW>(Bd + gi); },
        [&](int, int, const VD& b, bool) { return SLog<VW>{b, vsplat<VW>(1.f)}; });
  }
  double* Cd = (double*)a.Ch + hslot(a, s - 1) * a.stage + img;
  tile_inner_pass<VW>(
      T, tx, ty, [&](int) { return 0; },
      [&](int li, int gi, bool ok, int) {
        if (ok) vstd<VW>(Cd + gi, tile_log_result<VW>(sm.buf, sm.L, T, li).x);
      });
}

// log phi^s = K (log p - log c^{s-1}) (:704-707).
template <int VW>
__global__ __launch_bounds__(kThreads, 6) void k_tfwd2_log(SolveArgs a, int s) {
  typedef typename Vec<VW>::T V;
  typedef typename VecD<VW>::T VD;
  int p, i, tile;
  if (!map_tile(a, p, i, tile) || !fwd_runs(a, p, s)) return;
  extern __shared__ __align__(16) float smem[];
  const TileGeom T = a.T;
  const TileSmem sm = tile_smem(smem, T);
  load_tables(sm.tabs, a.tab);
  const int tx = tile % T.nx, ty = tile / T.nx;
  const long long N = a.N;
  const long long img = ((long long)p * a.K + i) * N;
  const double* Cs = (const double*)a.Ch + hslot(a, s - 1) * a.stage + img;
  const float* LA = a.logA + (long long)i * N;
  struct TwoL { VD c; V la; };
  tile_log_conv<VW>(
      sm.buf, sm.tabs, sm.L, T, tx, ty, a.events,
      [&](int gi) {
        TwoL r;
        r.c = vldd<VW>(Cs + gi);
        r.la = vld<VW>(LA + gi);
        return r;
      },
      [&](int, int, const TwoL& r, bool) { return SLog<VW>{vd(r.la) - r.c, vsplat<VW>(1.f)}; });
  double* Fd = (double*)a.Fh + hslot(a, s - 1) * a.stage + img;
  double lsum = 0;
  tile_inner_pass<VW>(
      T, tx, ty, [&](int) { return 0; },
      [&](int li, int gi, bool ok, int) {
        if (!ok) return;
        const VD f = tile_log_result<VW>(sm.buf, sm.L, T, li).x;
        vstd<VW>(Fd + gi, f);
        lsum += vsum_d(f);
      });
  const double t = block_sum(lsum);
  if (threadIdx.x == 0) a.Mt[mt_index(a, s, p, i, tile)] = t;
}

// Backward step s, first convolution, log domain.  chain 0 (:740-756): c = K
// [(lam_i n + v_i) b^s] with signs, resultDic += sign(c) exp(c - log c^{s-1});
// chain 1 (:778-796): the dot <log phi^s, g>, c = K [(lam_i g + r_i) / phi^s].
// Both store the second convolution's input log p + c - 2 log c^{s-1}, sign(c).
// Held to 8 waves per SIMD (64 VGPRs, a few bytes of spills): four workgroups
// per CU instead of three (55.2k -> 56.5k Sinkhorn-iters/s on the 256x256 log
// leg, 80-row regions).
template <int VW>
__global__ __launch_bounds__(kThreads, 8) void k_tbwd1_log(SolveArgs a, int t) {
  typedef typename Vec<VW>::T V;
  typedef typename VecD<VW>::T VD;
  int p, i, tile, chain;
  if (!map_tile_bwd(a, p, i, tile, chain)) return;
  const int s = bwd_step(a, p, t);
  if (s < 1) return;
  extern __shared__ __align__(16) float smem[];
  const TileGeom T = a.T;
  const TileSmem sm = tile_smem(smem, T);
  load_tables(sm.tabs, a.tab);
  const int tx = tile % T.nx, ty = tile / T.nx;
  const long long N = a.N;
  const int K = a.K;
  const long long hbase = (long long)p * K * N;
  const long long img = hbase + (long long)i * N;
  const float lpi = a.lam[(long long)p * K + i];
  const float* LA = a.logA + (long long)i * N;
  const double* Cs = (const double*)a.Ch + hslot(a, s - 1) * a.stage + img;  // log c^{s-1}
  const float* G0 = a.G0 + (long long)p * N;
  const float* NJ = a.nj + (long long)p * N;
  const float* GJ = a.gj + (long long)p * N;
  const bool last = (s == steps_of(a, p));
  struct Three { V x, y; VD z; };
  struct Mid { VD c; V la, gd; };
  // X = log p + c - 2 log c^{s-1}, sign of c (:754-756, :796)
  double* X = (double*)(chain == 0 ? a.X0 : a.X1) + img;
  unsigned char* XS = (chain == 0 ? a.XS0 : a.XS1) + img;
  auto mid_store = [&](int li, int gi, const Mid& r) {
    const SLog<VW> c = tile_log_result<VW>(sm.buf, sm.L, T, li);
    vstd<VW>(X + gi, vd(r.la) + c.x - 2.0 * r.c);
    SignIO<VW>::store(XS + gi, c.s);
    return c;
  };
  if (chain == 0) {
    const float* Vp = (s & 1) ? a.V0 : a.V1;
    const double* Bs = (const double*)a.Bh + hslot(a, s - 1) * a.stage + img;  // log b^s
    tile_log_conv<VW>(
        sm.buf, sm.tabs, sm.L, T, tx, ty, a.events,
        [&](int gi) {
          Three r;  // n, v_i, log b^s
          r.z = vldd<VW>(Bs + gi);
          if (last) {
            r.x = vld<VW>(G0 + gi);
            r.y = vsplat<VW>(0.f);
          } else {
            r.x = vld<VW>(NJ + gi);  // n = -sum_i v_i (:764), from k_tjoin
            r.y = vld<VW>(Vp + img + gi);
          }
          return r;
        },
        [&](int, int, const Three& r, bool second) {
          const V x = lpi * r.x + r.y;
          return SLog<VW>{vhlog_sel(vabs(x), second) + r.z, vsign(x)};
        });
    float* GDp = a.GD + img;
    tile_inner_pass<VW>(
        T, tx, ty,
        [&](int gi) {
          Mid r;
          r.c = vldd<VW>(Cs + gi);
          r.la = vld<VW>(LA + gi);
          r.gd = last ? vsplat<VW>(0.f) : vld<VW>(GDp + gi);
          return r;
        },
        [&](int li, int gi, bool ok, const Mid& r) {
          if (!ok) return;
          SLog<VW> c;
          if (s > 1) {
            c = mid_store(li, gi, r);
          } else {
            c = tile_log_result<VW>(sm.buf, sm.L, T, li);
          }
          vst<VW>(GDp + gi, r.gd + c.s * vexp_d(c.x - r.c));  // :752
        });
  } else {
    const float* Rp = (s & 1) ? a.R0 : a.R1;
    const double* Fs = (const double*)a.Fh + hslot(a, s - 1) * a.stage + img;  // log phi^s
    const float* Qp = a.Q + (long long)p * N;
    const double m = last ? 0.0 : mean_log_phi(a, s, p, i);
    double dot = 0;
    auto wload = [&](int gi) {
      Three r;  // g, r_i, log phi^s
      r.z = vldd<VW>(Fs + gi);
      if (last) {
        r.x = vld<VW>(G0 + gi) * vld<VW>(Qp + gi);
        r.y = vsplat<VW>(0.f);
      } else {
        r.x = vld<VW>(GJ + gi);  // g = -sum_i r_i (:808), from k_tjoin
        r.y = vld<VW>(Rp + img + gi);
      }
      return r;
    };
    auto acc_dot = [&](const Three& r) { dot += tdot(r.z, r.x) - m * vsum_d(r.x); };  // :786
    if (s == 1) {
      tile_inner_pass<VW>(T, tx, ty, wload, [&](int, int, bool ok, const Three& r) {
        if (ok) acc_dot(r);
      });
    } else {
      tile_log_conv<VW>(sm.buf, sm.tabs, sm.L, T, tx, ty, a.events, wload,
                        [&](int li, int, const Three& r, bool second) {
                          if (second && tile_inner(T, li)) acc_dot(r);
                          const V x = lpi * r.x + r.y;
                          return SLog<VW>{vhlog_sel(vabs(x), second) - r.z, vsign(x)};
                        });
    }
    double tot = block_sum(dot);
    if (threadIdx.x == 0) {
      if (!last && tile == 0) tot += m * exact_gsum(a, p, s);
      double* gl = a.GLt + ((long long)p * K + i) * (T.nx * T.ny) + tile;
      *gl = last ? tot : *gl + tot;
    }
    if (s == 1) return;
    tile_inner_pass<VW>(
        T, tx, ty,
        [&](int gi) {
          Mid r;
          r.c = vldd<VW>(Cs + gi);
          r.la = vld<VW>(LA + gi);
          r.gd = vsplat<VW>(0.f);
          return r;
        },
        [&](int li, int gi, bool ok, const Mid& r) {
          if (ok) mid_store(li, gi, r);
        });
  }
}

// Backward step s >= 2, second convolution, log domain: -v = sign exp(K X0 - log
// phi^{s-1}) (:757-762), -r = sign exp(K X1 + log b^{s-1}) (:797-805).
template <int VW>
__global__ __launch_bounds__(kThreads, 6) void k_tbwd2_log(SolveArgs a, int t) {
  typedef typename Vec<VW>::T V;
  typedef typename VecD<VW>::T VD;
  int p, i, tile, chain;
  if (!map_tile_bwd(a, p, i, tile, chain)) return;
  const int s = bwd_step(a, p, t);
  if (s < 2) return;
  extern __shared__ __align__(16) float smem[];
  const TileGeom T = a.T;
  const TileSmem sm = tile_smem(smem, T);
  load_tables(sm.tabs, a.tab);
  const int tx = tile % T.nx, ty = tile / T.nx;
  const long long N = a.N;
  const long long img = ((long long)p * a.K + i) * N;
  const double* X = (const double*)(chain == 0 ? a.X0 : a.X1) + img;
  const unsigned char* XS = (chain == 0 ? a.XS0 : a.XS1) + img;
  struct In { VD x; V s; };
  tile_log_conv<VW>(
      sm.buf, sm.tabs, sm.L, T, tx, ty, a.events,
      [&](int gi) {
        In r;
        r.x = vldd<VW>(X + gi);
        r.s = SignIO<VW>::load(XS + gi);
        return r;
      },
      [&](int, int, const In& r, bool) { return SLog<VW>{r.x, r.s}; });
  if (chain == 0) {
    const double* Fm = (const double*)a.Fh + hslot(a, s - 2) * a.stage + img;  // log phi^{s-1}
    float* Vc = ((s & 1) ? a.V1 : a.V0) + img;
    tile_inner_pass<VW>(
        T, tx, ty, [&](int gi) { return vldd<VW>(Fm + gi); },
        [&](int li, int gi, bool ok, VD f) {
          if (!ok) return;
          const SLog<VW> v = tile_log_result<VW>(sm.buf, sm.L, T, li);
          vst<VW>(Vc + gi, v.s * vexp_d(v.x - f));  // -v (:762)
        });
  } else {
    const double* Bm = (const double*)a.Bh + hslot(a, s - 2) * a.stage + img;  // log b^{s-1}
    float* Rc = ((s & 1) ? a.R1 : a.R0) + img;
    tile_inner_pass<VW>(
        T, tx, ty, [&](int gi) { return vldd<VW>(Bm + gi); },
        [&](int li, int gi, bool ok, VD b) {
          if (!ok) return;
          const SLog<VW> r = tile_log_result<VW>(sm.buf, sm.L, T, li);
          vst<VW>(Rc + gi, r.s * vexp_d(r.x + b));  // -r (:805)
        });
  }
}

// ---- launchers -------------------------------------------------------------------

int tiled_lds_bytes(const SolveArgs& a) {
  return tile_lds_bytes(a.T) + (a.logA ? tile_keys(a.T) * (int)sizeof(int) : 0);
}

int tiled_set_lds_attr() {
  static std::atomic<bool> done{false};  // solvers are created from several threads
  if (done.load(std::memory_order_relaxed)) return WDL_OK;
  TileGeom big{};
  big.R = kTile + 2 * kHaloMax;
  big.S = kMaxSide + 4;
  const int bytes = tile_lds_bytes(big) + kLogKeysBytes;
  const void* fns[] = {(const void*)k_tfwd1<1>,     (const void*)k_tfwd1<4>,
                       (const void*)k_tfwd2<1>,     (const void*)k_tfwd2<4>,
                       (const void*)k_tbwd1<1>,     (const void*)k_tbwd1<4>,
                       (const void*)k_tbwd2<1>,     (const void*)k_tbwd2<4>,
                       (const void*)k_tfwd1_log<1>, (const void*)k_tfwd1_log<4>,
                       (const void*)k_tfwd2_log<1>, (const void*)k_tfwd2_log<4>,
                       (const void*)k_tbwd1_log<1>, (const void*)k_tbwd1_log<4>,
                       (const void*)k_tbwd2_log<1>, (const void*)k_tbwd2_log<4>};
  for (const void* f : fns)
    WDL_HIP(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
  done.store(true, std::memory_order_relaxed);
  return WDL_OK;
}

#define WDL_TLAUNCH(kern, grid, ...)                                                     \
  do {                                                                                   \
    if (v4)                                                                              \
      hipLaunchKernelGGL(kern<4>, dim3(grid), dim3(kThreads), lds, st, __VA_ARGS__);     \
    else                                                                                 \
      hipLaunchKernelGGL(kern<1>, dim3(grid), dim3(kThreads), lds, st, __VA_ARGS__);     \
  } while (0)

int tiled_forward(const SolveArgs& a, hipStream_t st, int s) {
  const int lds = tiled_lds_bytes(a);
  const int grid = tile_grid(a);
  const bool v4 = a.T.W % 4 == 0;
  if (a.logA) {
    if (s >= 2) {
      const long long n = (long long)a.np * a.N;
      hipLaunchKernelGGL(k_tlogb, dim3((unsigned)std::min<long long>((n + 255) / 256, 8192)),
                         dim3(256), 0, st, a, s);
    }
    WDL_TLAUNCH(k_tfwd1_log, grid, a, s);
    WDL_TLAUNCH(k_tfwd2_log, grid, a, s);
  } else {
    if (s >= 2) {
      const int vw = v4 ? 4 : 1;
      const long long per = (a.N + 256LL * vw - 1) / (256LL * vw);
      const dim3 g((unsigned)std::min<long long>(per, 256), (unsigned)a.np);
      if (v4)
        hipLaunchKernelGGL(k_tq<4>, g, dim3(256), 0, st, a, s);
      else
        hipLaunchKernelGGL(k_tq<1>, g, dim3(256), 0, st, a, s);
    }
    WDL_TLAUNCH(k_tfwd1, grid, a, s);
    WDL_TLAUNCH(k_tfwd2, grid, a, s);
  }
  WDL_HIP(hipGetLastError());
  return WDL_OK;
}

int tiled_backward(const SolveArgs& a, hipStream_t st, int t) {
  const int lds = tiled_lds_bytes(a);
  const int grid = (a.weightsOnly ? 1 : 2) * tile_grid(a);
  const bool v4 = a.T.W % 4 == 0;
  if (a.logA) {
    const long long n = (long long)a.np * a.N;
    hipLaunchKernelGGL(k_tjoin, dim3((unsigned)std::min<long long>((n + 255) / 256, 8192)),
                       dim3(256), 0, st, a, t);
    WDL_TLAUNCH(k_tbwd1_log, grid, a, t);
    if (t >= 2) WDL_TLAUNCH(k_tbwd2_log, grid, a, t);
  } else {
    WDL_TLAUNCH(k_tbwd1, grid, a, t);
    if (t >= 2) WDL_TLAUNCH(k_tbwd2, grid, a, t);
  }
  WDL_HIP(hipGetLastError());
  return WDL_OK;
}

int tiled_reduce_gl(const SolveArgs& a, hipStream_t st) {
  const int n = a.np * a.K;
  hipLaunchKernelGGL(k_tgl, dim3((n + 255) / 256), dim3(256), 0, st, a);
  WDL_HIP(hipGetLastError());
  return WDL_OK;
}

#undef WDL_TLAUNCH

}  // namespace wdl
