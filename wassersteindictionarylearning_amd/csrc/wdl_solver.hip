// Batched Sinkhorn barycenter + unrolled gradient, standard domain
// (compute_gradient_both, inverseWasserstein.h:544-674), for P histograms at
// once, and the evaluate_both callback around it (:986-1145).
//
// Device layout (FP32 unless noted; pixel index fastest):
//   A   [K][N]        atoms (normalised)            lam [P][K]   obs [P][N]
//   Qh  [L][P][N]     Qh[s-1] = q^s   (s = 1..L); the history b^s = q^s / phi^s
//                     (:587-592) is recomputed from it bit for bit
//   (log domain: Bh [L][P][K][N] FP64, log b^s)
//   Ch  [L][P][K][N]  Ch[s-1] = c^{s-1} = K b^{s-1}  history conv_b
//   Fh  [L][P][K][N]  Fh[s-1] = phi^s = K a^s        history phi
//   V0/V1, R0/R1 [P][K][N]  ping-pong v (atom chain) and r (weight chain); during
//                     the forward V0/V1 carry safe_log phi^s between stages
//   GD  [P][K][N]  resultDic per histogram;  GL [P][K] (FP64) resultLa
//   Q, G0 [P][N]   barycenter q^L and the loss gradient at q^L
//
// Each Sinkhorn stage is one launch over (histogram, atom) workgroups; a
// workgroup keeps its image in LDS through two convolutions with the
// elementwise steps fused around them.  Cross-atom reductions (the geometric
// mean, n = sum_i v_i, g = sum_i r_i) are recomputed by each of the K
// workgroups of a histogram from the previous stage's output; the K
// workgroups of one histogram share an XCD so these re-reads hit its L2.
#include <atomic>
#include <cmath>
#include <cstring>
#include <limits>

#include "wdl_solve_common.h"

namespace wdl {

// Forward iteration s (1..L): b^{s-1} (geometric-mean step of iteration s-1,
// :577-592), c^{s-1} = K b^{s-1} (:568), a = p / c (:569-574), phi^s = K a (:575).
//
// Workgroup shape of a stage: THREADS threads, two workgroups per CU (the
// image's LDS allows no more).  FP32 passes: 768 or 1024 threads for the
// forward (its launch has one workgroup per (histogram, atom), fewer than
// the CUs hold, so each image can take more waves), 512 or 768 for the
// backward, picked per solver by the tuner (default 1024 / 768); the
// split-BF16x3 passes need more registers and keep 512.
template <int THREADS> struct StageCfg {
  static constexpr int threads = THREADS;
  static constexpr int waves = THREADS / 64;
  static constexpr int waves_per_simd = THREADS / 128;  // 2 workgroups over 4 SIMDs
};
template <int VW, bool BF3, int THREADS>
__global__ __launch_bounds__(THREADS, StageCfg<THREADS>::waves_per_simd) void k_forward(
    SolveArgs a, int s) {
  constexpr int kFwdThreads = THREADS, kFwdWaves = StageCfg<THREADS>::waves;
  typedef typename Vec<VW>::T V;
  int p, i;
  if (!map_block(a, p, i) || !fwd_runs(a, p, s)) return;
  extern __shared__ __align__(16) float smem[];
  void* tabs = smem;
  float* buf = smem + ConvPrec<BF3>::kTabBytes / 4;
  ConvPrec<BF3>::template load<kFwdThreads>(tabs, a.tab);
  const ConvGeom G = a.G;
  const long long N = a.N;
  const int K = a.K;
  const long long hbase = (long long)p * K * N;
  const long long img = hbase + (long long)i * N;
  if (s == 1) {
    const float* bi = a.bInit ? a.bInit + img : nullptr;
    px_pass<VW, kUnrollPx, kFwdThreads>(
        G, [&](int gi) { return bi ? vld<VW>(bi + gi) : vsplat<VW>(1.f); },
        [&](int li, int gi, bool ok, V b) { vst<VW>(buf + li, ok ? b : vsplat<VW>(0.f)); });
  } else {
    // safe_log phi_j^{s-1} of every atom j, written by stage s-1 (each
    // workgroup logs its own phi once instead of all K of them)
    const float* Lf = (((s - 1) & 1) ? a.LF1 : a.LF0) + hbase;
    const float* Fi = a.Fh + hslot(a, s - 2) * a.stage + img;  // phi_i^{s-1}
    float* Qd = i == 0 ? a.Qh + hslot(a, s - 2) * a.qstage + (long long)p * N : nullptr;
    const float* lam = a.lam + (long long)p * K;
    struct LdF { V lq, fi; };
    px_pass<VW, kUnrollPx, kFwdThreads>(
        G,
        [&](int gi) {
          LdF r;
          r.lq = vsplat<VW>(0.f);
          for_atoms<VW>(Lf + gi, N, K, [&](int j, V v) { r.lq += lam[j] * v; });
          r.fi = vld<VW>(Fi + gi);
          return r;
        },
        [&](int li, int gi, bool ok, const LdF& r) {
          V b = vsplat<VW>(0.f);
          if (ok) {
            const V q = vexpf(r.lq);
            if (Qd) vst<VW>(Qd + gi, q);
            b = fdiv(q, r.fi);
          }
          vst<VW>(buf + li, b);
        });
  }
  ConvPrec<BF3>::template conv<kFwdWaves>(buf, tabs, G);
  {
    float* Cd = a.Ch + hslot(a, s - 1) * a.stage + img;
    const float* Ai = a.A + (long long)i * N;
    px_pass<VW, kUnrollPx, kFwdThreads>(
        G, [&](int gi) { return vld<VW>(Ai + gi); },
        [&](int li, int gi, bool ok, V av) {
          V u = vsplat<VW>(0.f);
          if (ok) {
            const V c = vld<VW>(buf + li);
            vst<VW>(Cd + gi, c);
            u = fdiv(av, c);
          }
          vst<VW>(buf + li, u);
        });
  }
  ConvPrec<BF3>::template conv<kFwdWaves>(buf, tabs, G);
  float* Fd = a.Fh + hslot(a, s - 1) * a.stage + img;
  // the next stage's geometric mean reads safe_log phi^s (LF0 / LF1 are the
  // V buffers, free until the backward writes them, unless histories are
  // checkpointed; k_final logs phi^L itself)
  float* Ld = s < steps_of(a, p) ? ((s & 1) ? a.LF1 : a.LF0) + img : nullptr;
  px_pass<VW, kUnrollPx, kFwdThreads>(
      G, [&](int gi) { return 0; },
      [&](int li, int gi, bool ok, int) {
        if (!ok) return;
        const V f = vld<VW>(buf + li);
        vst<VW>(Fd + gi, f);
        if (Ld) vst<VW>(Ld + gi, safe_logf(f));
      });
}

// After iteration L: q = exp(sum_i lam_i safe_log phi_i^L), b^L = q / phi^L,
// the loss and its gradient at q (:577-597, :1089).
__global__ __launch_bounds__(kThreads) void k_final(SolveArgs a) {
  const int pl = blockIdx.y;
  const int p = a.p0 + pl;
  const long long N = a.N;
  const int K = a.K;
  const long long hbase = (long long)p * K * N;
  const int Lp = steps_of(a, p);
  if (Lp == 0) {  // idle histogram (continuous batching): no loss, no gradient
    if (threadIdx.x == 0) {
      a.lossPart[(long long)p * kLossBlocks + blockIdx.x] = 0.0;
      a.gsumPart[(long long)p * kLossBlocks + blockIdx.x] = 0.0;
    }
    return;
  }
  const float* F = a.Fh + hslot(a, Lp - 1) * a.stage + hbase;
  float* QL = a.Qh + hslot(a, Lp - 1) * a.qstage + (long long)p * N;
  const float* lam = a.lam + (long long)p * K;
  double acc = 0, gs = 0, bad = 0;
  for (long long x = (long long)blockIdx.x * kThreads + threadIdx.x; x < N;
       x += (long long)gridDim.x * kThreads) {
    float lq = 0.f;
    bool ok = true;
    for (int j = 0; j < K; ++j) {
      const float f = F[j * N + x];
      ok = ok && f > 0.f && f < INFINITY;
      lq += lam[j] * safe_logf(f);
    }
    const float q = expf(lq);
    ok = ok && q > 0.f && q < INFINITY;
    if (!ok) bad += 1.0;  // the FP32 scalings left their range at this pixel
    QL[x] = q;
    if (a.bSave)
      for (int j = 0; j < K; ++j) a.bSave[hbase + j * N + x] = fdiv(q, F[j * N + x]);
    a.Q[(long long)p * N + x] = q;
    if (a.loss == WDL_LOSS_WASSERSTEIN) continue;  // set by wasserstein_loss
    const float o = a.obs[(long long)p * N + x];
    const float g0 = loss_grad(a.loss, q, o);
    a.G0[(long long)p * N + x] = g0;
    acc += loss_term(a.loss, q, o);
    gs += (double)(g0 * q);  // g_L as the weight chain forms it
  }
  const double t = block_sum(acc), tg = block_sum(gs), tb = block_sum(bad);
  if (threadIdx.x == 0 && a.loss != WDL_LOSS_WASSERSTEIN) {
    a.lossPart[(long long)p * kLossBlocks + blockIdx.x] = t;
    a.gsumPart[(long long)p * kLossBlocks + blockIdx.x] = tg;
  }
  if (threadIdx.x == 0 && tb > 0 && a.events) atomicAdd(a.events, (unsigned)tb);
}

// Backward step s (L..1).  chain 0: dictionary gradient (:600-629);
// chain 1: weight gradient (:635-673).
//
// Workgroup shape as the forward's (StageCfg): 512 or 768 threads for the
// FP32 passes, 512 for split-BF16x3.
template <int VW, bool BF3, int THREADS>
__global__ __launch_bounds__(THREADS, StageCfg<THREADS>::waves_per_simd) void k_backward(
    SolveArgs a, int t) {
  constexpr int kBwdThreads = THREADS, kBwdWaves = StageCfg<THREADS>::waves;
  typedef typename Vec<VW>::T V;
  int p, i, chain;
  if (!map_block_bwd(a, p, i, chain)) return;
  const int s = bwd_step(a, p, t);
  if (s < 1) return;
  extern __shared__ __align__(16) float smem[];
  void* tabs = smem;
  float* buf = smem + ConvPrec<BF3>::kTabBytes / 4;
  ConvPrec<BF3>::template load<kBwdThreads>(tabs, a.tab);
  const ConvGeom G = a.G;
  const long long N = a.N;
  const int K = a.K;
  const long long hbase = (long long)p * K * N;
  const long long img = hbase + (long long)i * N;
  const float lpi = a.lam[(long long)p * K + i];
  const float* Ai = a.A + (long long)i * N;
  const float* Cs = a.Ch + hslot(a, s - 1) * a.stage + img;  // c^{s-1}
  const bool last = (s == steps_of(a, p));
  struct Two { V x, y; };
  struct Three { V x, y, z; };
  if (chain == 0) {
    const float* Vp = (s & 1) ? a.V0 : a.V1;
    float* Vc = (s & 1) ? a.V1 : a.V0;
    const float* Qs = a.Qh + hslot(a, s - 1) * a.qstage + (long long)p * N;  // q^s
    const float* Fs = a.Fh + hslot(a, s - 1) * a.stage + img;                // phi^s
    const float* G0 = a.G0 + (long long)p * N;
    px_pass<VW, kUnrollPx, kBwdThreads>(
        G,
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
        [&](int li, int gi, bool ok, const Three& r) {
          vst<VW>(buf + li, ok ? (lpi * r.x - r.y) * r.z : vsplat<VW>(0.f));
        });
    ConvPrec<BF3>::template conv<kBwdWaves>(buf, tabs, G);
    float* GDp = a.GD + img;
    px_pass<VW, kUnrollPx, kBwdThreads>(
        G,
        [&](int gi) {
          Three r;  // conv_b, resultDic so far, atom
          r.x = vld<VW>(Cs + gi);
          r.y = last ? vsplat<VW>(0.f) : vld<VW>(GDp + gi);
          r.z = vld<VW>(Ai + gi);
          return r;
        },
        [&](int li, int gi, bool ok, const Three& r) {
          V t = vsplat<VW>(0.f);
          if (ok) {
            const V c = vld<VW>(buf + li);
            const V cq = fdiv(c, r.x);
            vst<VW>(GDp + gi, r.y + cq);
            t = fdiv(-r.z * cq, r.x);
          }
          vst<VW>(buf + li, t);
        });
    if (s == 1) return;
    ConvPrec<BF3>::template conv<kBwdWaves>(buf, tabs, G);
    const float* Fm = a.Fh + hslot(a, s - 2) * a.stage + img;  // phi^{s-1}
    px_pass<VW, kUnrollPx, kBwdThreads>(
        G, [&](int gi) { return vld<VW>(Fm + gi); },
        [&](int li, int gi, bool ok, V f) {
          if (ok) vst<VW>(Vc + img + gi, fdiv(vld<VW>(buf + li), f));
        });
  } else {
    const float* Rp = (s & 1) ? a.R0 : a.R1;
    float* Rc = (s & 1) ? a.R1 : a.R0;
    const float* Fs = a.Fh + hslot(a, s - 1) * a.stage + img;  // phi^s
    const float* G0 = a.G0 + (long long)p * N;
    const float* Qp = a.Q + (long long)p * N;
    double dot = 0, sf = 0, sg = 0;
    px_pass<VW, kUnrollPx, kBwdThreads>(
        G,
        [&](int gi) {
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
        },
        [&](int li, int gi, bool ok, const Three& r) {
          V u = vsplat<VW>(0.f);
          if (ok) {
            const V lf = vlogf(r.z);  // plain log (:649)
            dot += vdot_d(lf, r.x);
            sf += vsum_d(lf);
            sg += vsum_d(r.x);
            u = fdiv(lpi * r.x - r.y, r.z);
          }
          vst<VW>(buf + li, u);
        });
    const double t = block_sum<kBwdWaves>(dot), tf = block_sum<kBwdWaves>(sf), tg = block_sum<kBwdWaves>(sg);
    if (threadIdx.x == 0) {
      double* gl = a.GL + (long long)p * K + i;
      *gl = last ? t : *gl + (t - tf / (double)N * (tg - exact_gsum(a, p, s)));
    }
    if (s == 1) return;
    ConvPrec<BF3>::template conv<kBwdWaves>(buf, tabs, G);
    px_pass<VW, kUnrollPx, kBwdThreads>(
        G,
        [&](int gi) {
          Two r;
          r.x = vld<VW>(Cs + gi);
          r.y = vld<VW>(Ai + gi);
          return r;
        },
        [&](int li, int gi, bool ok, const Two& r) {
          vst<VW>(buf + li, ok ? vld<VW>(buf + li) * fdiv(fdiv(-r.y, r.x), r.x) : vsplat<VW>(0.f));
        });
    ConvPrec<BF3>::template conv<kBwdWaves>(buf, tabs, G);
    const float* Qm = a.Qh + hslot(a, s - 2) * a.qstage + (long long)p * N;  // q^{s-1}
    const float* Fm = a.Fh + hslot(a, s - 2) * a.stage + img;                // phi^{s-1}
    px_pass<VW, kUnrollPx, kBwdThreads>(
        G, [&](int gi) { return fdiv(vld<VW>(Qm + gi), vld<VW>(Fm + gi)); },  // b^{s-1}
        [&](int li, int gi, bool ok, V b) {
          if (ok) vst<VW>(Rc + img + gi, vld<VW>(buf + li) * b);
        });
  }
}

// GDsum[i][x] = sum_p GD[p][i][x] in FP64 (the per-histogram projection of
// :1103-1108 is linear, so it is applied to the sum on the host).
__global__ void k_reduce_gd(const float* __restrict__ GD, double* __restrict__ out, int P,
                            long long KN) {
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < KN;
       e += (long long)gridDim.x * blockDim.x) {
    double acc = 0;
    for (int p = 0; p < P; ++p) acc += GD[(long long)p * KN + e];
    out[e] = acc;
  }
}

__global__ void k_log(const float* __restrict__ x, float* __restrict__ y, long long n) {
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += (long long)gridDim.x * blockDim.x)
    y[e] = logf(x[e]);
}

__global__ void k_zero_u32(unsigned* p) { *p = 0u; }

// Checkpointed histories (WDL_FLAG_CHECKPOINT).  k_copy_ck saves phi^c of a
// group's histograms (n values from a history slot) at a segment boundary c.
template <class T>
__global__ void k_copy_ck(const T* __restrict__ src, T* __restrict__ dst, long long n) {
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += (long long)gridDim.x * blockDim.x)
    dst[e] = src[e];
}

// Before the segment after boundary c is recomputed: phi^c back into its
// history slot for the histograms whose backward reaches past the segment
// (L_p >= rmin; the others' slots hold entries their own segments still read),
// and in the resident standard domain its safe_log into the ping-pong the
// forward's geometric mean reads (the value k_forward wrote there).
template <class T>
__global__ void k_restore(SolveArgs a, const T* __restrict__ ck, T* __restrict__ fslot,
                          float* __restrict__ lf) {
  const long long KN = (long long)a.K * a.N, tot = (long long)a.np * KN;
  const long long o0 = (long long)a.p0 * KN;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < tot;
       e += (long long)gridDim.x * blockDim.x) {
    if (steps_of(a, a.p0 + (int)(e / KN)) < a.rmin) continue;
    const T v = ck[o0 + e];
    fslot[o0 + e] = v;
    if (lf) lf[o0 + e] = safe_logf((float)v);
  }
}

template <class T>
__global__ void k_fill(T* __restrict__ dst, T v, long long n) {
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += (long long)gridDim.x * blockDim.x)
    dst[e] = v;
}

// ---- log domain (compute_gradient_both_log_signArray, :678-814) ----------------
// Bh/Ch/Fh hold log b, log conv_b, log phi in FP64 (a log value needs
// absolute precision; FP32 would give it ~|x| * 6e-8); V/R hold the linear
// -v / -r the reference carries (:762, :805) and GD is linear, in FP32.  Each convolution is
// log_conv_lds (per-line max shifts + MFMA); its input is rebuilt from
// global data in both of its passes.

__device__ inline double vdot_dd(double a, float b) { return a * (double)b; }
__device__ inline double vdot_dd(doublex4 a, floatx4 b) {
  return a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w;
}

struct LogSmem {
  float* tabs;
  float* buf;
  LogLds L;
};

__device__ inline LogSmem log_smem(float* smem, const ConvGeom& G) {
  LogSmem m;
  m.tabs = smem;
  m.buf = smem + kTabFloats;
  const int img = (G.Hp * G.S1 > G.Wp * G.S2) ? G.Hp * G.S1 : G.Wp * G.S2;
  int* keys = reinterpret_cast<int*>(m.buf + img);
  m.L.mrow = keys;
  m.L.mcol = keys + kMaxSide;
  m.L.rmin = keys + 2 * kMaxSide;
  return m;
}

template <int VW>
__global__ __launch_bounds__(kThreads, 4) void k_forward_log(SolveArgs a, int s) {
  typedef typename Vec<VW>::T V;
  int p, i;
  if (!map_block(a, p, i) || !fwd_runs(a, p, s)) return;
  extern __shared__ __align__(16) float smem[];
  const ConvGeom G = a.G;
  const LogSmem sm = log_smem(smem, G);
  load_tables(sm.tabs, a.tab);
  const long long N = a.N;
  const int K = a.K;
  const long long hbase = (long long)p * K * N;
  const long long img = hbase + (long long)i * N;
  typedef typename VecD<VW>::T VD;
  struct TwoD { VD x, y; };
  // conv 1 input: log b^{s-1} (b^{s-1} = lq - phi_i, :711-724)
  if (s == 1) {
    const double* bi = a.lbInit ? a.lbInit + img : nullptr;  // warm start (log b, FP64)
    log_conv_lds<VW>(
        sm.buf, sm.tabs, sm.L, G, a.events,
        [&](int gi) { return bi ? vldd<VW>(bi + gi) : vd(vsplat<VW>(0.f)); },
        [&](int, int, int, VD b, bool) { return SLog<VW>{b, vsplat<VW>(1.f)}; });
  } else {
    const double* Fp = (const double*)a.Fh + hslot(a, s - 2) * a.stage + hbase;  // log phi^{s-1}
    double* Bd = (double*)a.Bh + hslot(a, s - 2) * a.stage + img;
    const float* lam = a.lam + (long long)p * K;
    log_conv_lds<VW>(
        sm.buf, sm.tabs, sm.L, G, a.events,
        [&](int gi) {
          TwoD r;
          r.x = vd(vsplat<VW>(0.f));
          r.y = vd(vsplat<VW>(0.f));
          for_atoms_l<4>(
              K, [&](int j) { return vldd<VW>(Fp + (long long)j * N + gi); },
              [&](int j, const VD& f) {
                r.x += (double)lam[j] * f;
                if (j == i) r.y = f;
              });
          return r;
        },
        [&](int, int gi, int, const TwoD& r, bool second) {
          const VD b = r.x - r.y;
          if (second) vstd<VW>(Bd + gi, b);
          return SLog<VW>{b, vsplat<VW>(1.f)};
        });
  }
  // conv 2 input: log a = log p - log c (:704-707); store log c
  double* Cd = (double*)a.Ch + hslot(a, s - 1) * a.stage + img;
  const float* LA = a.logA + (long long)i * N;
  log_conv_lds<VW>(
      sm.buf, sm.tabs, sm.L, G, a.events, [&](int gi) { return vld<VW>(LA + gi); },
      [&](int li, int gi, int row, V la, bool second) {
        const SLog<VW> c = log_result<VW>(sm.buf, sm.L, G, li, row);
        if (second) vstd<VW>(Cd + gi, c.x);
        return SLog<VW>{vd(la) - c.x, vsplat<VW>(1.f)};
      });
  double* Fd = (double*)a.Fh + hslot(a, s - 1) * a.stage + img;
  px_pass_r<VW>(
      G, [&](int) { return 0; },
      [&](int li, int gi, int row, bool ok, int) {
        if (ok) vstd<VW>(Fd + gi, log_result<VW>(sm.buf, sm.L, G, li, row).x);
      });
}

// After iteration L: log q = sum_i lam_i log phi_i^L, b^L = log q - log phi^L,
// q = exp(log q) (:711-728), loss and its gradient at q.
// bx / nbx: the workgroup's pixel chunk and the chunks per histogram; by: the
// histogram (from p0).
__device__ inline void final_log_body(const SolveArgs& a, int bx, int nbx, int by) {
  const int p = a.p0 + by;
  const long long N = a.N;
  const int K = a.K;
  const long long hbase = (long long)p * K * N;
  const int Lp = steps_of(a, p);
  if (Lp == 0) {
    if (threadIdx.x == 0) {
      a.lossPart[(long long)p * kLossBlocks + bx] = 0.0;
      a.gsumPart[(long long)p * kLossBlocks + bx] = 0.0;
    }
    return;
  }
  const double* F = (const double*)a.Fh + hslot(a, Lp - 1) * a.stage + hbase;
  double* B = (double*)a.Bh + hslot(a, Lp - 1) * a.stage + hbase;
  const float* lam = a.lam + (long long)p * K;
  double acc = 0, gs = 0;
  for (long long x = (long long)bx * kThreads + threadIdx.x; x < N;
       x += (long long)nbx * kThreads) {
    double lq = 0.0;
    for (int j = 0; j < K; ++j) lq += (double)lam[j] * F[j * N + x];
    for (int j = 0; j < K; ++j) {
      const double b = lq - F[j * N + x];
      B[j * N + x] = b;
      if (a.lbSave) a.lbSave[hbase + j * N + x] = b;
    }
    const float q = (float)exp(lq);
    a.Q[(long long)p * N + x] = q;
    if (a.loss == WDL_LOSS_WASSERSTEIN) continue;  // set by wasserstein_loss
    const float o = a.obs[(long long)p * N + x];
    const float g0 = loss_grad(a.loss, q, o);
    a.G0[(long long)p * N + x] = g0;
    acc += loss_term(a.loss, q, o);
    gs += (double)(g0 * q);
  }
  const double t = block_sum(acc), tg = block_sum(gs);
  if (threadIdx.x == 0 && a.loss != WDL_LOSS_WASSERSTEIN) {
    a.lossPart[(long long)p * kLossBlocks + bx] = t;
    a.gsumPart[(long long)p * kLossBlocks + bx] = tg;
  }
}

__global__ __launch_bounds__(kThreads) void k_final_log(SolveArgs a) {
  final_log_body(a, blockIdx.x, gridDim.x, blockIdx.y);
}

// A group's final stage: kLossBlocks workgroups per histogram of each member.
__global__ __launch_bounds__(kThreads) void k_final_log_group(GroupArgs g) {
  int lb, nb;
  const SolveArgs a = g.m[group_member(g, GK_FINAL, blockIdx.x, lb, nb)].a;
  final_log_body(a, lb % kLossBlocks, kLossBlocks, lb / kLossBlocks);
}

template <int VW>
__global__ __launch_bounds__(kThreads, 4) void k_backward_log(SolveArgs a, int t) {
  typedef typename Vec<VW>::T V;
  int p, i, chain;
  if (!map_block_bwd(a, p, i, chain)) return;
  const int s = bwd_step(a, p, t);
  if (s < 1) return;
  extern __shared__ __align__(16) float smem[];
  const ConvGeom G = a.G;
  const LogSmem sm = log_smem(smem, G);
  load_tables(sm.tabs, a.tab);
  const long long N = a.N;
  const int K = a.K;
  const long long hbase = (long long)p * K * N;
  const long long img = hbase + (long long)i * N;
  const float lpi = a.lam[(long long)p * K + i];
  const float* LA = a.logA + (long long)i * N;
  const double* Cs = (const double*)a.Ch + hslot(a, s - 1) * a.stage + img;  // log c^{s-1}
  const bool last = (s == steps_of(a, p));
  typedef typename VecD<VW>::T VD;
  struct Three { V x, y; VD z; };
  struct Two { V x, y; };
  struct TwoL { VD x; V y; };
  // second conv input: log p + c - 2 log c^{s-1}, with the sign of c (:754-756, :796)
  auto mid_make = [&](int li, int, int row, const TwoL& r, bool) {
    const SLog<VW> c = log_result<VW>(sm.buf, sm.L, G, li, row);
    return SLog<VW>{vd(r.y) + c.x - 2.0 * r.x, c.s};
  };
  auto mid_load = [&](int gi) {
    TwoL r;
    r.x = vldd<VW>(Cs + gi);
    r.y = vld<VW>(LA + gi);
    return r;
  };
  if (chain == 0) {  // dictionary gradient, :736-770
    const float* Vp = (s & 1) ? a.V0 : a.V1;
    float* Vc = (s & 1) ? a.V1 : a.V0;
    const double* Bs = (const double*)a.Bh + hslot(a, s - 1) * a.stage + img;  // log b^s
    const float* G0 = a.G0 + (long long)p * N;
    log_conv_lds<VW>(
        sm.buf, sm.tabs, sm.L, G, a.events,
        [&](int gi) {
          Three r;  // n, v_i, log b^s
          r.z = vldd<VW>(Bs + gi);
          if (last) {
            r.x = vld<VW>(G0 + gi);
            r.y = vsplat<VW>(0.f);
          } else {
            r.x = vsplat<VW>(0.f);
            r.y = vsplat<VW>(0.f);
            for_atoms<VW>(Vp + hbase + gi, N, K, [&](int j, V v) {
              r.x -= v;  // n = -sum_i v_i (:764)
              if (j == i) r.y = v;
            });
          }
          return r;
        },
        [&](int, int, int, const Three& r, bool second) {
          const V x = lpi * r.x + r.y;
          return SLog<VW>{vhlog_sel(vabs(x), second) + r.z, vsign(x)};
        });
    float* GDp = a.GD + img;
    px_pass_r<VW>(
        G,
        [&](int gi) {
          TwoL r;
          r.x = vldd<VW>(Cs + gi);
          r.y = last ? vsplat<VW>(0.f) : vld<VW>(GDp + gi);
          return r;
        },
        [&](int li, int gi, int row, bool ok, const TwoL& r) {
          if (ok) {
            const SLog<VW> c = log_result<VW>(sm.buf, sm.L, G, li, row);
            vst<VW>(GDp + gi, r.y + c.s * vexpf(vf(c.x - r.x)));  // :752
          }
        });
    if (s == 1) return;
    log_conv_lds<VW>(sm.buf, sm.tabs, sm.L, G, a.events, mid_load, mid_make);
    const double* Fm = (const double*)a.Fh + hslot(a, s - 2) * a.stage + img;  // log phi^{s-1}
    px_pass_r<VW>(
        G, [&](int gi) { return vldd<VW>(Fm + gi); },
        [&](int li, int gi, int row, bool ok, VD f) {
          if (ok) {
            const SLog<VW> v = log_result<VW>(sm.buf, sm.L, G, li, row);
            vst<VW>(Vc + img + gi, v.s * vexpf(vf(v.x - f)));  // -v (:762)
          }
        });
  } else {  // weight gradient, :774-811
    const float* Rp = (s & 1) ? a.R0 : a.R1;
    float* Rc = (s & 1) ? a.R1 : a.R0;
    const double* Fs = (const double*)a.Fh + hslot(a, s - 1) * a.stage + img;  // log phi^s
    const float* G0 = a.G0 + (long long)p * N;
    const float* Qp = a.Q + (long long)p * N;
    double dot = 0, sf = 0, sg = 0;
    auto wload = [&](int gi) {
      Three r;  // g, r_i, log phi^s
      r.z = vldd<VW>(Fs + gi);
      if (last) {
        r.x = vld<VW>(G0 + gi) * vld<VW>(Qp + gi);
        r.y = vsplat<VW>(0.f);
      } else {
        r.x = vsplat<VW>(0.f);
        r.y = vsplat<VW>(0.f);
        for_atoms<VW>(Rp + hbase + gi, N, K, [&](int j, V v) {
          r.x -= v;  // g = -sum_i r_i (:808)
          if (j == i) r.y = v;
        });
      }
      return r;
    };
    auto acc_dot = [&](const Three& r) {
      dot += vdot_dd(r.z, r.x);  // sum phi . g (:786)
      sf += vsum_d(r.z);
      sg += vsum_d(r.x);
    };
    auto wmake = [&](int, int, int, const Three& r, bool second) {
      if (second) acc_dot(r);
      const V x = lpi * r.x + r.y;
      return SLog<VW>{vhlog_sel(vabs(x), second) - r.z, vsign(x)};
    };
    if (s == 1) {
      px_pass_r<VW>(G, wload, [&](int, int, int, bool ok, const Three& r) {
        if (ok) acc_dot(r);
      });
    } else {
      log_conv_lds<VW>(sm.buf, sm.tabs, sm.L, G, a.events, wload, wmake);
    }
    const double t = block_sum(dot), tf = block_sum(sf), tg = block_sum(sg);
    if (threadIdx.x == 0) {
      double* gl = a.GL + (long long)p * K + i;
      *gl = last ? t : *gl + (t - tf / (double)N * (tg - exact_gsum(a, p, s)));
    }
    if (s == 1) return;
    log_conv_lds<VW>(sm.buf, sm.tabs, sm.L, G, a.events, mid_load, mid_make);
    const double* Bm = (const double*)a.Bh + hslot(a, s - 2) * a.stage + img;  // log b^{s-1}
    px_pass_r<VW>(
        G, [&](int gi) { return vldd<VW>(Bm + gi); },
        [&](int li, int gi, int row, bool ok, VD b) {
          if (ok) {
            const SLog<VW> r = log_result<VW>(sm.buf, sm.L, G, li, row);
            vst<VW>(Rc + img + gi, r.s * vexpf(vf(r.x + b)));  // -r (:805)
          }
        });
  }
}

// ---- projection rounds (continuous batching) ------------------------------------
// One mirror-descent step on the simplex per active histogram: with the
// weight gradient G = resultLa (:635-673) and its simplex-tangent projection
// g = G - <lambda, G> (as evaluate_both, :1095-1099),
//   lambda_i <- lambda_i exp(-eta g_i / max_j |g_j|) / Z,
// the exponentiated-gradient update with a step normalised by the largest
// component (scale-free across losses and grid sizes).
__device__ inline void project_update_one(float* __restrict__ lam, const double* __restrict__ GL,
                                          const int* __restrict__ Lp,
                                          const float* __restrict__ eta, int p, int P, int K) {
  if (p >= P || Lp[p] <= 0) return;
  float* l = lam + (long long)p * K;
  const double* g = GL + (long long)p * K;
  double d = 0;
  for (int i = 0; i < K; ++i) d += (double)l[i] * g[i];
  double m = 0;
  for (int i = 0; i < K; ++i) m = fmax(m, fabs(g[i] - d));
  if (!(m > 0) || !isfinite(m)) return;
  const double step = (double)eta[p] / m;
  double z = 0;
  for (int i = 0; i < K; ++i) z += (double)l[i] * exp(-step * (g[i] - d));
  for (int i = 0; i < K; ++i) l[i] = (float)((double)l[i] * exp(-step * (g[i] - d)) / z);
}

__global__ void k_project_update(float* __restrict__ lam, const double* __restrict__ GL,
                                 const int* __restrict__ Lp, const float* __restrict__ eta, int P,
                                 int K) {
  project_update_one(lam, GL, Lp, eta, blockIdx.x * blockDim.x + threadIdx.x, P, K);
}

// b_storage <- b_temp for the active histograms (the warm restart of :418-422,
// per histogram); bx / nbx: the workgroup's share of histogram p's K N values.
template <class T>
__device__ inline void warm_commit_one(T* __restrict__ store, const T* __restrict__ temp,
                                       const int* __restrict__ Lp, long long KN, int p, int bx,
                                       int nbx) {
  if (Lp[p] <= 0) return;
  const long long o = (long long)p * KN;
  for (long long e = (long long)bx * blockDim.x + threadIdx.x; e < KN;
       e += (long long)nbx * blockDim.x)
    store[o + e] = temp[o + e];
}

template <class T>
__global__ void k_warm_commit_active(T* __restrict__ store, const T* __restrict__ temp,
                                     const int* __restrict__ Lp, long long KN) {
  warm_commit_one(store, temp, Lp, KN, blockIdx.y, blockIdx.x, gridDim.x);
}

// A group's projection round after its solve: the update (64 histograms per
// workgroup) and the warm commit (min(64, K N / 256) workgroups per histogram)
// of every member; members are log-domain, so the store holds log b in FP64.
__global__ void k_project_update_group(GroupArgs g) {
  int lb, nb;
  const GroupMember& gm = g.m[group_member(g, GK_PROJ, blockIdx.x, lb, nb)];
  project_update_one(const_cast<float*>(gm.a.lam), gm.a.GL, gm.a.Lp, gm.eta,
                     lb * blockDim.x + threadIdx.x, gm.a.P, gm.a.K);
}

__global__ void k_warm_commit_group(GroupArgs g) {
  int lb, nb;
  const SolveArgs a = g.m[group_member(g, GK_WARM, blockIdx.x, lb, nb)].a;
  const long long KN = (long long)a.K * a.N;
  const int per = nb / a.P;  // workgroups per histogram
  warm_commit_one(const_cast<double*>(a.lbInit), (const double*)a.lbSave, a.Lp, KN, lb / per,
                  lb % per, per);
}

// A group's summed atom gradients (full-gradient members), as k_reduce_gd.
__global__ void k_reduce_gd_group(GroupArgs g) {
  int lb, nb;
  const GroupMember& gm = g.m[group_member(g, GK_RGD, blockIdx.x, lb, nb)];
  const float* GD = gm.a.GD;
  double* out = gm.GDsum;
  const int P = gm.a.P;
  const long long KN = (long long)gm.a.K * gm.a.N;
  for (long long e = (long long)lb * blockDim.x + threadIdx.x; e < KN;
       e += (long long)nb * blockDim.x) {
    double acc = 0;
    for (int p = 0; p < P; ++p) acc += GD[(long long)p * KN + e];
    out[e] = acc;
  }
}

// Vector width 4 when rows hold whole float4 groups; the workgroup shape of
// the FP32 stages is a.fwdT / a.bwdT.
template <int VW, bool BF3, int THREADS>
static void fwd_launch(const SolveArgs& a, int grid, int lds, hipStream_t st, int s) {
  hipLaunchKernelGGL((k_forward<VW, BF3, THREADS>), dim3(grid), dim3(THREADS), lds, st, a, s);
}
template <int VW, bool BF3, int THREADS>
static void bwd_launch(const SolveArgs& a, int grid, int lds, hipStream_t st, int s) {
  hipLaunchKernelGGL((k_backward<VW, BF3, THREADS>), dim3(grid), dim3(THREADS), lds, st, a, s);
}

static void launch_forward(const SolveArgs& a, int grid, int lds, hipStream_t st, int s) {
  if (a.logA) {
    if (a.G.W % 4 == 0)
      hipLaunchKernelGGL(k_forward_log<4>, dim3(grid), dim3(kThreads), lds, st, a, s);
    else
      hipLaunchKernelGGL(k_forward_log<1>, dim3(grid), dim3(kThreads), lds, st, a, s);
    return;
  }
  const bool v4 = a.G.W % 4 == 0;
  if (a.bf3) {
    if (v4) fwd_launch<4, true, kThreads>(a, grid, lds, st, s);
    else fwd_launch<1, true, kThreads>(a, grid, lds, st, s);
  } else if (a.fwdT == 768) {
    if (v4) fwd_launch<4, false, 768>(a, grid, lds, st, s);
    else fwd_launch<1, false, 768>(a, grid, lds, st, s);
  } else {
    if (v4) fwd_launch<4, false, 1024>(a, grid, lds, st, s);
    else fwd_launch<1, false, 1024>(a, grid, lds, st, s);
  }
}

static void launch_backward(const SolveArgs& a, int grid1, int lds, hipStream_t st, int s) {
  const int blocks = a.weightsOnly ? grid1 : 2 * grid1;  // one or two chains per (p, i)
  if (a.logA) {
    if (a.G.W % 4 == 0)
      hipLaunchKernelGGL(k_backward_log<4>, dim3(blocks), dim3(kThreads), lds, st, a, s);
    else
      hipLaunchKernelGGL(k_backward_log<1>, dim3(blocks), dim3(kThreads), lds, st, a, s);
    return;
  }
  const bool v4 = a.G.W % 4 == 0;
  if (a.bf3) {
    if (v4) bwd_launch<4, true, kThreads>(a, blocks, lds, st, s);
    else bwd_launch<1, true, kThreads>(a, blocks, lds, st, s);
  } else if (a.bwdT == 512) {
    if (v4) bwd_launch<4, false, 512>(a, blocks, lds, st, s);
    else bwd_launch<1, false, 512>(a, blocks, lds, st, s);
  } else {
    if (v4) bwd_launch<4, false, 768>(a, blocks, lds, st, s);
    else bwd_launch<1, false, 768>(a, blocks, lds, st, s);
  }
}

}  // namespace wdl

using namespace wdl;

struct wdl_solver {
  wdl_ctx* ctx = nullptr;
  int W = 0, H = 0, K = 0, P = 0, L = 0, loss = 0, flags = 0;
  double gamma = 0;
  long long N = 0;
  int D = 0;
  DevBuf tab, A, lam, obs, Bh, Qh, Ch, Fh, V0, V1, R0, R1, GD, GL, Q, G0, lossPart, bStore,
      bTemp, GDsum, logA, events, Lp, X0, X1, XS0, XS1, GLt, gsumPart, Mt, eta, xa, xt, xdp,
      nj, gj, sga, sgt, wla, wlb, wlc, wlt, evX, evG, evW, evA, evF, evP, CK, LF0, LF1, lmax;
  std::vector<int> hostLp;  // the per-histogram iteration counts; lmax holds their maximum
  int seg = 0, ring = 0;  // checkpointed histories: segment length, history slots (0: L slots)
  int fwdT = 1024, bwdT = 768, xwpl = 0, xlw = 4;  // launch shape (wdl_solver_set_tuning)
  bool tiled = false;  // 64 x 64 tiles with a halo (a side above kMaxSide)
  bool exact = false;  // log domain with the exact-shift line passes (WDL_FLAG_EXACT_LSE)
  TileGeom tg{};
  // The batched run as two captured graphs: [0] the first forward steps
  // (none when L < 8), [1] the rest.
  hipGraph_t graph[2] = {nullptr, nullptr};
  hipGraphExec_t exec[2] = {nullptr, nullptr};
  bool captured = false;
  // Steps enqueue() emits: its launch sequence is numbered (forward steps,
  // the loss, backward steps or segments, the reductions) and only numbers in
  // [op_lo, op_hi) are launched.
  int op_lo = 0, op_hi = std::numeric_limits<int>::max();
  bool observed = false, problem = false, ran = false;
  double scaleDic = 1.0;
  int expWeight = 1;
  std::vector<double> hostAtoms, hostLam;  // normalised FP64 copies

  SolveArgs args(int p0, int np) const {
    SolveArgs a;
    a.G = make_geom(W, H, (flags & WDL_FLAG_DENSE) ? (1 << 20) : D,
                    (flags & WDL_FLAG_BF16X3) ? 32 : 16);
    a.bf3 = (flags & WDL_FLAG_BF16X3) ? 1 : 0;
    a.K = K;
    a.P = P;
    a.L = L;
    a.p0 = p0;
    a.np = np;
    a.loss = loss;
    a.N = N;
    a.stage = (long long)P * K * N;
    a.tab = tab.as<float>();
    a.A = A.as<float>();
    a.lam = lam.as<float>();
    a.obs = obs.as<float>();
    a.Bh = Bh.as<float>();
    a.Qh = Qh.as<float>();
    a.qstage = (long long)P * N;
    a.Ch = Ch.as<float>();
    a.Fh = Fh.as<float>();
    a.V0 = V0.as<float>();
    a.V1 = V1.as<float>();
    a.R0 = R0.as<float>();
    a.R1 = R1.as<float>();
    a.GD = GD.as<float>();
    a.GL = GL.as<double>();
    a.Q = Q.as<float>();
    a.G0 = G0.as<float>();
    a.lossPart = lossPart.as<double>();
    const bool warm = flags & WDL_FLAG_WARM_RESTART;
    const bool logw = warm && (flags & WDL_FLAG_LOG_DOMAIN);
    a.bInit = warm && !logw ? bStore.as<float>() : nullptr;
    a.bSave = warm && !logw ? bTemp.as<float>() : nullptr;
    a.lbInit = logw ? bStore.as<double>() : nullptr;
    a.lbSave = logw ? bTemp.as<double>() : nullptr;
    a.Lp = Lp.as<int>();
    a.ring = ring;
    a.rmin = 0;
    a.fwdT = fwdT;
    a.bwdT = bwdT;
    a.xwpl = xwpl;
    a.xlw = xlw;
    a.LF0 = ring && !(flags & WDL_FLAG_LOG_DOMAIN) ? LF0.as<float>() : V0.as<float>();
    a.LF1 = ring && !(flags & WDL_FLAG_LOG_DOMAIN) ? LF1.as<float>() : V1.as<float>();
    a.weightsOnly = (flags & WDL_FLAG_WEIGHTS_ONLY) ? 1 : 0;
    const bool logd = flags & WDL_FLAG_LOG_DOMAIN;
    a.logA = logd ? logA.as<float>() : nullptr;
    a.events = events.as<unsigned>();
    a.T = tg;
    a.X0 = X0.as<float>();
    a.X1 = X1.as<float>();
    a.XS0 = XS0.as<unsigned char>();
    a.XS1 = XS1.as<unsigned char>();
    a.GLt = GLt.as<double>();
    a.gsumPart = gsumPart.as<double>();
    a.Mt = Mt.as<double>();
    a.gamma = gamma;
    a.wla = wla.as<double>();
    a.wlb = wlb.as<double>();
    a.wlc = wlc.as<double>();
    a.wlt = wlt.as<double>();
    a.xa = xa.as<double>();
    a.xt = xt.as<double>();
    a.xdp = xdp.as<double>();
    a.xfwd = (flags & WDL_FLAG_WEIGHTS_ONLY) ? 0 : 2 * a.stage;
    a.nj = nj.as<float>();
    a.gj = gj.as<float>();
    a.sga = sga.as<unsigned char>();
    a.sgt = sgt.as<unsigned char>();
    return a;
  }

  int lds_bytes(const SolveArgs& a) const {
    if (tiled) return tiled_lds_bytes(a);
    return conv_lds_bytes(a.G) + ((flags & WDL_FLAG_LOG_DOMAIN) ? kLogKeysBytes : 0) +
           (a.bf3 ? ConvPrec<true>::kTabBytes - ConvPrec<false>::kTabBytes : 0);
  }

  // Side streams for the batched form: histogram groups advance on their own
  // stream, so stages of different groups overlap on the CUs (one group's
  // elementwise HBM passes run beside another group's MFMA passes).
  int ngroups = 1;
  std::vector<hipStream_t> side;
  std::vector<hipEvent_t> joins;
  hipEvent_t fork = nullptr;

  // The captured graph holds the launch shape; a new shape captures again.
  int drop_graph() {
    WDL_HIP(hipStreamSynchronize(ctx->stream));
    release_graphs();
    return WDL_OK;
  }
  void release_graphs() {
    for (int k = 0; k < 2; ++k) {
      if (exec[k]) (void)hipGraphExecDestroy(exec[k]);
      if (graph[k]) (void)hipGraphDestroy(graph[k]);
      exec[k] = nullptr;
      graph[k] = nullptr;
    }
    captured = false;
  }

  int set_groups(int n) {
    n = std::max(1, std::min(n, P));
    if (n == ngroups) return WDL_OK;
    int st;
    if ((st = drop_graph())) return st;
    release_streams();
    ngroups = n;
    if (n == 1) return WDL_OK;
    WDL_HIP(hipEventCreateWithFlags(&fork, hipEventDisableTiming));
    for (int g = 0; g < n; ++g) {
      hipStream_t st;
      hipEvent_t ev;
      WDL_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
      side.push_back(st);
      WDL_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
      joins.push_back(ev);
    }
    return WDL_OK;
  }

  void release_streams() {
    for (hipStream_t st : side) (void)hipStreamDestroy(st);
    for (hipEvent_t ev : joins) (void)hipEventDestroy(ev);
    if (fork) (void)hipEventDestroy(fork);
    side.clear();
    joins.clear();
    fork = nullptr;
  }

  // All P solves; with several groups, fork onto the side streams and join
  // back on the context stream before the final reduction.
  int enqueue_batched() {
    if (op_lo == 0)
      hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(1), 0, ctx->stream, events.as<unsigned>());
    if (ngroups <= 1) return enqueue(0, P, true, ctx->stream);
    WDL_HIP(hipEventRecord(fork, ctx->stream));
    int st;
    for (int g = 0; g < ngroups; ++g) {
      const int p0 = (int)((long long)P * g / ngroups), p1 = (int)((long long)P * (g + 1) / ngroups);
      WDL_HIP(hipStreamWaitEvent(side[g], fork, 0));
      if ((st = enqueue(p0, p1 - p0, false, side[g]))) return st;
      WDL_HIP(hipEventRecord(joins[g], side[g]));
      WDL_HIP(hipStreamWaitEvent(ctx->stream, joins[g], 0));
    }
    const bool last = op_hi == std::numeric_limits<int>::max();
    return (!last || (flags & WDL_FLAG_WEIGHTS_ONLY)) ? WDL_OK : reduce_gd(ctx->stream);
  }

  int reduce_gd(hipStream_t st) {
    const long long KN = (long long)K * N;
    const int g = (int)std::min<long long>(1024, (KN + 255) / 256);
    hipLaunchKernelGGL(k_reduce_gd, dim3(g), dim3(256), 0, st, GD.as<float>(), GDsum.as<double>(),
                       P, KN);
    WDL_HIP(hipGetLastError());
    return WDL_OK;
  }

  int enqueue(int p0, int np, bool reduce, hipStream_t st) {
    const SolveArgs a = args(p0, np);
    const int lds = lds_bytes(a);
    const dim3 blk(kThreads);
    const int gs = stage_grid(np, K);
    int e;
    auto forward = [&](const SolveArgs& x, int s) -> int {
      if (exact) return exact_forward(x, st, s);
      if (tiled) return tiled_forward(x, st, s);
      launch_forward(x, gs, lds, st, s);
      return WDL_OK;
    };
    auto backward = [&](int s) -> int {
      if (exact) return exact_backward(a, st, s);
      if (tiled) return tiled_backward(a, st, s);
      launch_backward(a, gs, lds, st, s);
      return WDL_OK;
    };
    const bool logd = flags & WDL_FLAG_LOG_DOMAIN;
    const size_t esz = logd ? sizeof(double) : sizeof(float);
    const long long gofs = (long long)p0 * K * N, gn = (long long)np * K * N;  // the group's values
    const dim3 cgrid((unsigned)std::min<long long>((gn + 255) / 256, 4096));
    int op = 0;
    auto take = [&]() {
      const int o = op++;
      return o >= op_lo && o < op_hi;
    };
    for (int s = 1; s <= L; ++s) {
      if (!take()) continue;
      if ((e = forward(a, s))) return e;
      if (ring && s % seg == 0 && s < L) {  // checkpoint phi^s
        char* src = (char*)Fh.p + (size_t)(hslot(a, s - 1) * a.stage + gofs) * esz;
        char* dst = (char*)CK.p + (size_t)((long long)(s / seg - 1) * a.stage + gofs) * esz;
        if (logd)
          hipLaunchKernelGGL(k_copy_ck<double>, cgrid, dim3(256), 0, st, (const double*)src,
                             (double*)dst, gn);
        else
          hipLaunchKernelGGL(k_copy_ck<float>, cgrid, dim3(256), 0, st, (const float*)src,
                             (float*)dst, gn);
      }
    }
    // kLossBlocks blocks per histogram: every lossPart slot is written
    // (blocks past the last pixel write 0).
    if (take()) {
      if (a.logA)
        hipLaunchKernelGGL(k_final_log, dim3(kLossBlocks, np), blk, 0, st, a);
      else
        hipLaunchKernelGGL(k_final, dim3(kLossBlocks, np), blk, 0, st, a);
      if (loss == WDL_LOSS_WASSERSTEIN && (e = wasserstein_loss(a, st))) return e;
    }
    if (!ring) {
      for (int s = L; s >= 1; --s)
        if (take() && (e = backward(s))) return e;
    } else {
      // Segments (c, c + seg] from the last: the ring still holds the last
      // segment of every histogram after the forward (a histogram's last
      // ring entries are its own last iterations); an earlier segment is
      // recomputed from its checkpoint for the histograms whose backward
      // reaches past it, one iteration beyond it so q^e / log b^e exist.
      const int nseg = (L + seg - 1) / seg;
      for (int k = nseg - 1; k >= 0; --k) {
        const int c = k * seg, end = std::min(c + seg, L);
        if (end < L && take()) {
          SolveArgs r = a;
          r.rmin = end + 1;
          r.events = nullptr;  // the first pass counted these lines
          if (c > 0) {
            const long long ck = (long long)(k - 1) * a.stage, fs = hslot(a, c - 1) * a.stage;
            float* lf = (!logd && !tiled && !exact) ? ((c & 1) ? r.LF1 : r.LF0) : nullptr;
            const dim3 rg((unsigned)std::min<long long>((gn + 255) / 256, 4096));
            if (logd)
              hipLaunchKernelGGL(k_restore<double>, rg, dim3(256), 0, st, r,
                                 CK.as<double>() + ck, Fh.as<double>() + fs, lf);
            else
              hipLaunchKernelGGL(k_restore<float>, rg, dim3(256), 0, st, r,
                                 CK.as<float>() + ck, Fh.as<float>() + fs, lf);
          }
          for (int s = c + 1; s <= end + 1; ++s)
            if ((e = forward(r, s))) return e;
        }
        for (int s = end; s > c; --s)
          if (take() && (e = backward(s))) return e;
      }
    }
    if (!take()) return WDL_OK;
    if (tiled && !exact && (e = tiled_reduce_gl(a, st))) return e;
    WDL_HIP(hipGetLastError());
    return (reduce && !(flags & WDL_FLAG_WEIGHTS_ONLY)) ? reduce_gd(st) : WDL_OK;
  }
};

static int set_lds_attr() {
  static std::atomic<bool> done{false};  // solvers are created from several threads
  if (done.load(std::memory_order_relaxed)) return WDL_OK;
  const int bytes = conv_lds_bytes(make_geom(kMaxSide, kMaxSide, 0));
  const void* fns[] = {
      (const void*)k_forward<1, false, 768>,   (const void*)k_forward<4, false, 768>,
      (const void*)k_forward<1, false, 1024>,  (const void*)k_forward<4, false, 1024>,
      (const void*)k_backward<1, false, 512>,  (const void*)k_backward<4, false, 512>,
      (const void*)k_backward<1, false, 768>,  (const void*)k_backward<4, false, 768>};
  for (const void* f : fns)
    WDL_HIP(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
  const void* bfns[] = {(const void*)k_forward<1, true, kThreads>,
                        (const void*)k_forward<4, true, kThreads>,
                        (const void*)k_backward<1, true, kThreads>,
                        (const void*)k_backward<4, true, kThreads>};
  for (const void* f : bfns)
    WDL_HIP(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize,
                                bytes + ConvPrec<true>::kTabBytes - ConvPrec<false>::kTabBytes));
  const void* lfns[] = {(const void*)k_forward_log<1>, (const void*)k_forward_log<4>,
                        (const void*)k_backward_log<1>, (const void*)k_backward_log<4>};
  for (const void* f : lfns)
    WDL_HIP(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize,
                                bytes + kLogKeysBytes));
  done.store(true, std::memory_order_relaxed);
  return WDL_OK;
}

static bool use_exact(int flags) {
  return (flags & WDL_FLAG_EXACT_LSE) && (flags & WDL_FLAG_LOG_DOMAIN);
}

static bool use_tiles(int W, int H, int flags) {
  return !use_exact(flags) && (needs_tiles(W, H) || (flags & WDL_FLAG_TILED));
}

// Checkpointed histories: segments of seg = ceil(sqrt(L)) iterations, a ring
// of seg + 2 history slots (a segment's entries, the one before it and the one
// its recomputation writes beyond it) and one checkpoint of phi per segment
// boundary.  ring = 0 (L slots) when that would not take less memory: the
// ring's three histories, the checkpoints and two ping-pong images against
// three histories of L slots.
static void ckpt_shape(int L, int flags, int& seg, int& ring, int& nck) {
  seg = ring = nck = 0;
  if (!(flags & WDL_FLAG_CHECKPOINT)) return;
  const int sg = (int)std::ceil(std::sqrt((double)L));
  if (3 * (sg + 2) + (L + sg - 1) / sg - 1 + 2 >= 3 * L) return;
  seg = sg;
  ring = sg + 2;
  nck = (L + sg - 1) / sg - 1;
}

static long long footprint(int W, int H, int K, int P, int L, int flags) {
  const long long N = (long long)W * H, PKN = (long long)P * K * N;
  const bool logd = flags & WDL_FLAG_LOG_DOMAIN;
  int seg, ring, nck;
  ckpt_shape(L, flags, seg, ring, nck);
  const long long hs = ring ? ring : L;  // history slots
  long long f = (flags & WDL_FLAG_LOG_DOMAIN) ? 3LL * hs * PKN * 8           // Bh, Ch, Fh (FP64)
                                             : 2LL * hs * PKN * 4 + hs * P * N * 4;  // Ch, Fh, Qh
  if (ring) f += (long long)nck * PKN * (logd ? 8 : 4) + (logd ? 0 : 2 * PKN * 4);  // CK, LF0/1
  f += 5 * PKN * 4;                          // V0, V1, R0, R1, GD
  f += (long long)K * N * (4 + 8);           // A, GDsum
  f += (long long)P * N * 4 * 3;             // obs, Q, G0
  f += (long long)P * K * (4 + 8) + 2LL * P * kLossBlocks * 8 + (long long)P * 4;
  if (flags & WDL_FLAG_WARM_RESTART) f += 2 * PKN * (logd ? 8 : 4);  // log domain: log b, FP64
  if (flags & WDL_FLAG_LOG_DOMAIN) f += (long long)K * N * 4;  // log atoms
  f += kTabLen * 4;
  f += 4LL * P * N * 8;  // the Wasserstein loss's log scalings and convolution buffers
  f += (3LL * P * K + 3LL * K * N + 1) * 8;  // evaluate_both: x, g, weights, atoms (FP64), f
  f += 2LL * K * ev_chunks(N) * 8;            // and the atoms' per-chunk sums
  if (use_exact(flags)) {  // FP64 scratch images xa, xt + sign bytes, the dots' partial sums
    const long long nchb = (flags & WDL_FLAG_WEIGHTS_ONLY) ? 1 : 2, xim = nchb == 1 ? 1 : 3;
    f += 2LL * PKN * (xim * 8 + nchb);
    f += (long long)P * K * xdot_chunks((long long)W * H) * 3 * 8;
  }
  if (use_tiles(W, H, flags)) {  // X0, X1 (+ sign bytes), per-tile dots, log: cross-atom sums
    const long long nt = (long long)((W + kTile - 1) / kTile) * ((H + kTile - 1) / kTile);
    f += 2 * PKN * (logd ? 9 : 4) + (long long)P * K * nt * 8 * (1 + L);
    if (logd) f += 2LL * P * N * 4;
  }
  return f;
}

extern "C" {

long long wdl_solver_footprint(int W, int H, int K, int P, int L, int flags) {
  if (W <= 0 || H <= 0 || K <= 0 || P <= 0 || L <= 0) return -1;
  return footprint(W, H, K, P, L, flags);
}

static std::atomic<int> g_debug_fill{0};

int wdl_debug_fill(int on) { return g_debug_fill.exchange(on ? 1 : 0); }

int wdl_solver_create(wdl_ctx* ctx, int W, int H, int K, int P, int L, double gamma, int loss,
                      int flags, wdl_solver** out) {
  WDL_CHECK(ctx && out, WDL_ERR_ARG, "wdl_solver_create: NULL argument");
  WDL_CHECK(W > 0 && H > 0 && K > 0 && P > 0 && L > 0, WDL_ERR_ARG,
            "wdl_solver_create: W, H, K, P, L must be positive");
  WDL_CHECK(gamma > 0 && std::isfinite(gamma), WDL_ERR_ARG,
            "wdl_solver_create: gamma must be positive and finite");
  WDL_CHECK(loss == WDL_LOSS_QUADRATIC || loss == WDL_LOSS_KL || loss == WDL_LOSS_TV ||
                loss == WDL_LOSS_WASSERSTEIN,
            WDL_ERR_UNSUPPORTED, "wdl_solver_create: loss must be 1, 2, 3 or 4 (main_dictionary_learning.cpp:189-197)");
  WDL_CHECK(loss != WDL_LOSS_WASSERSTEIN || std::max(W, H) <= kXMaxLen, WDL_ERR_UNSUPPORTED,
            "wdl_solver_create: the Wasserstein loss takes sides up to " +
                std::to_string(kXMaxLen));
  WDL_CHECK(!(flags & WDL_FLAG_BF16X3) ||
                (!(flags & (WDL_FLAG_LOG_DOMAIN | WDL_FLAG_TILED)) && !needs_tiles(W, H)),
            WDL_ERR_UNSUPPORTED,
            "wdl_solver_create: WDL_FLAG_BF16X3 is for the standard domain with sides <= " +
                std::to_string(kMaxSide));
  WDL_CHECK(!(flags & WDL_FLAG_EXACT_LSE) || (flags & WDL_FLAG_LOG_DOMAIN), WDL_ERR_ARG,
            "wdl_solver_create: WDL_FLAG_EXACT_LSE needs WDL_FLAG_LOG_DOMAIN");
  WDL_CHECK(!use_exact(flags) || std::max(W, H) <= kXMaxLen, WDL_ERR_UNSUPPORTED,
            "wdl_solver_create: exact-shift passes take sides up to " + std::to_string(kXMaxLen));
  const bool tiled = use_tiles(W, H, flags);
  WDL_CHECK(!(tiled && (flags & WDL_FLAG_DENSE)), WDL_ERR_UNSUPPORTED,
            "wdl_solver_create: WDL_FLAG_DENSE needs sides of at most " + std::to_string(kMaxSide) +
                " and no WDL_FLAG_TILED");
  std::vector<float> tab;
  const int D = build_table(W, H, gamma, tab);
  TileGeom tg{};
  int st;
  if (tiled && (st = make_tile_geom(W, H, D, gamma, flags & WDL_FLAG_LOG_DOMAIN, tg))) return st;
  WDL_HIP(hipSetDevice(ctx->device));
  if ((st = set_lds_attr())) return st;
  if ((st = tiled_set_lds_attr())) return st;
  if ((use_exact(flags) || loss == WDL_LOSS_WASSERSTEIN) &&
      (st = xlse_set_lds_attr(std::max(W, H))))
    return st;
  wdl_solver* s = new wdl_solver();
  s->tiled = tiled;
  s->exact = use_exact(flags);
  s->tg = tg;
  s->ctx = ctx;
  s->W = W;
  s->H = H;
  s->K = K;
  s->P = P;
  s->L = L;
  s->loss = loss;
  s->flags = flags;
  s->gamma = gamma;
  s->N = (long long)W * H;
  s->D = D;
  const long long N = s->N, PKN = (long long)P * K * N;
  const long long ntiles = (long long)tg.nx * tg.ny;
  const size_t f4 = sizeof(float);
  const size_t hsz = (flags & WDL_FLAG_LOG_DOMAIN) ? sizeof(double) : sizeof(float);  // histories
  int nck;
  ckpt_shape(L, flags, s->seg, s->ring, nck);
  const size_t hs = s->ring ? (size_t)s->ring : (size_t)L;  // history slots
  // exact-shift scratch images per (histogram, atom): the backward's chains,
  // plus the forward's own when there are two (SolveArgs::xfwd)
  const size_t nchb = (flags & WDL_FLAG_WEIGHTS_ONLY) ? 1 : 2, xim = nchb == 1 ? 1 : 3;
  struct {
    DevBuf* b;
    size_t n;
  } allocs[] = {
      {&s->tab, tab.size() * f4},
      {&s->A, (size_t)K * N * f4},
      {&s->lam, (size_t)P * K * f4},
      {&s->obs, (size_t)P * N * f4},
      {&s->Bh, (flags & WDL_FLAG_LOG_DOMAIN) ? (size_t)hs * PKN * hsz : 0},
      {&s->Qh, (flags & WDL_FLAG_LOG_DOMAIN) ? 0 : (size_t)hs * P * N * f4},
      {&s->Ch, (size_t)hs * PKN * hsz},
      {&s->Fh, (size_t)hs * PKN * hsz},
      {&s->CK, (size_t)nck * PKN * hsz},
      {&s->LF0, s->ring && !(flags & WDL_FLAG_LOG_DOMAIN) ? (size_t)PKN * f4 : 0},
      {&s->LF1, s->ring && !(flags & WDL_FLAG_LOG_DOMAIN) ? (size_t)PKN * f4 : 0},
      {&s->V0, (size_t)PKN * f4},
      {&s->V1, (size_t)PKN * f4},
      {&s->R0, (size_t)PKN * f4},
      {&s->R1, (size_t)PKN * f4},
      {&s->GD, (size_t)PKN * f4},
      {&s->GL, (size_t)P * K * sizeof(double)},
      {&s->Q, (size_t)P * N * f4},
      {&s->G0, (size_t)P * N * f4},
      {&s->lossPart, (size_t)P * kLossBlocks * sizeof(double)},
      {&s->GDsum, (size_t)K * N * sizeof(double)},
      {&s->bStore, (flags & WDL_FLAG_WARM_RESTART) ? (size_t)PKN * hsz : 0},
      {&s->bTemp, (flags & WDL_FLAG_WARM_RESTART) ? (size_t)PKN * hsz : 0},
      {&s->logA, (flags & WDL_FLAG_LOG_DOMAIN) ? (size_t)K * N * f4 : 0},
      {&s->events, sizeof(unsigned)},
      {&s->Lp, (size_t)P * sizeof(int)},
      {&s->lmax, sizeof(int)},
      {&s->X0, tiled ? (size_t)PKN * hsz : 0},
      {&s->X1, tiled ? (size_t)PKN * hsz : 0},
      {&s->XS0, tiled && (flags & WDL_FLAG_LOG_DOMAIN) ? (size_t)PKN : 0},
      {&s->XS1, tiled && (flags & WDL_FLAG_LOG_DOMAIN) ? (size_t)PKN : 0},
      {&s->nj, tiled && (flags & WDL_FLAG_LOG_DOMAIN) ? (size_t)P * N * f4 : 0},
      {&s->gj, tiled && (flags & WDL_FLAG_LOG_DOMAIN) ? (size_t)P * N * f4 : 0},
      {&s->GLt, tiled ? (size_t)P * K * ntiles * sizeof(double) : 0},
      {&s->Mt, tiled ? (size_t)L * P * K * ntiles * sizeof(double) : 0},
      {&s->gsumPart, (size_t)P * kLossBlocks * sizeof(double)},
      {&s->eta, (size_t)P * sizeof(float)},
      {&s->xa, use_exact(flags) ? (size_t)xim * PKN * sizeof(double) : 0},
      {&s->xt, use_exact(flags) ? (size_t)xim * PKN * sizeof(double) : 0},
      {&s->xdp, use_exact(flags) ? (size_t)P * K * xdot_chunks(N) * 3 * sizeof(double) : 0},
      {&s->sga, use_exact(flags) ? (size_t)nchb * PKN : 0},
      {&s->sgt, use_exact(flags) ? (size_t)nchb * PKN : 0},
      {&s->wla, (size_t)P * N * sizeof(double)},
      {&s->wlb, (size_t)P * N * sizeof(double)},
      {&s->wlc, (size_t)P * N * sizeof(double)},
      {&s->wlt, (size_t)P * N * sizeof(double)},
      {&s->evX, (size_t)(P * K + K * N) * sizeof(double)},
      {&s->evG, (size_t)(P * K + K * N) * sizeof(double)},
      {&s->evW, (size_t)P * K * sizeof(double)},
      {&s->evA, (size_t)K * N * sizeof(double)},
      {&s->evF, sizeof(double)},
      {&s->evP, (size_t)2 * K * ev_chunks(N) * sizeof(double)},
  };
  for (auto& al : allocs) {
    if ((st = al.b->ensure(al.n))) {
      wdl_solver_destroy(s);
      return st;
    }
  }
  if (g_debug_fill.load()) {  // wdl_debug_fill: NaN bytes until written
    for (auto& al : allocs)
      if (al.b->bytes && hipMemset(al.b->p, 0xFF, al.b->bytes) != hipSuccess) {
        wdl_solver_destroy(s);
        set_error("wdl_solver_create: debug fill failed");
        return WDL_ERR_HIP;
      }
  }
  hipError_t e = hipMemcpy(s->tab.p, tab.data(), tab.size() * f4, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    s->hostLp.assign((size_t)P, L);
    e = hipMemcpy(s->Lp.p, s->hostLp.data(), (size_t)P * sizeof(int), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(s->lmax.p, &L, sizeof(int), hipMemcpyHostToDevice);
  }
  // Outputs start at zero, so a histogram left idle (0 iterations) reads 0
  // until it first runs, and adds 0 to the summed atom gradient; so do the
  // mirror-descent steps, so a projection round over slots whose step was
  // never set only renormalises their weights.
  for (DevBuf* b : {&s->GD, &s->GL, &s->Q, &s->G0, &s->lossPart, &s->gsumPart, &s->GDsum,
                    &s->bTemp, &s->eta})
    if (e == hipSuccess && b->bytes) e = hipMemsetAsync(b->p, 0, b->bytes, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e == hipSuccess && (flags & WDL_FLAG_WARM_RESTART)) {
    // b_storage starts at 1 (inverseWasserstein.h:380-384).
    if (flags & WDL_FLAG_LOG_DOMAIN)
      hipLaunchKernelGGL(k_fill<double>, dim3(1024), dim3(256), 0, ctx->stream,
                         s->bStore.as<double>(), 1.0, PKN);
    else
      hipLaunchKernelGGL(k_fill<float>, dim3(1024), dim3(256), 0, ctx->stream,
                         s->bStore.as<float>(), 1.f, PKN);
    e = hipStreamSynchronize(ctx->stream);
  }
  if (e != hipSuccess) {
    wdl_solver_destroy(s);
    set_error(std::string("wdl_solver_create: ") + hipGetErrorString(e));
    return WDL_ERR_HIP;
  }
  // Default: one stream per 16 histograms, at most 4 (the hardware queues a
  // process gets by default).
  if ((st = s->set_groups(std::min(4, std::max(1, P / 16))))) {
    wdl_solver_destroy(s);
    return st;
  }
  *out = s;
  return WDL_OK;
}

int wdl_solver_set_streams(wdl_solver* s, int nstreams) {
  WDL_CHECK(s, WDL_ERR_ARG, "wdl_solver_set_streams: NULL argument");
  WDL_CHECK(nstreams >= 1 && nstreams <= 16, WDL_ERR_ARG,
            "wdl_solver_set_streams: nstreams must be in 1..16");
  WDL_HIP(hipSetDevice(s->ctx->device));
  return s->set_groups(nstreams);
}

int wdl_solver_destroy(wdl_solver* s) {
  if (!s) return WDL_OK;
  (void)hipSetDevice(s->ctx->device);
  (void)hipStreamSynchronize(s->ctx->stream);
  s->release_graphs();
  s->release_streams();
  DevBuf* bufs[] = {&s->tab, &s->A,  &s->lam, &s->obs, &s->Bh, &s->Qh, &s->Ch,  &s->Fh,
                    &s->V0,  &s->V1, &s->R0,  &s->R1,  &s->GD, &s->GL,       &s->Q,
                    &s->G0,  &s->lossPart,    &s->bStore,      &s->bTemp,    &s->GDsum,
                    &s->logA, &s->events, &s->Lp, &s->X0, &s->X1, &s->XS0, &s->XS1, &s->GLt, &s->Mt, &s->gsumPart, &s->eta, &s->xa, &s->xt, &s->xdp, &s->nj, &s->gj, &s->sga,
                    &s->sgt, &s->wla, &s->wlb, &s->wlc,
                    &s->wlt, &s->evX, &s->evG, &s->evW, &s->evA, &s->evF, &s->evP,
                    &s->CK, &s->LF0, &s->LF1, &s->lmax};
  for (DevBuf* b : bufs) b->release();
  delete s;
  return WDL_OK;
}

int wdl_solver_set_observed(wdl_solver* s, const double* observed) {
  WDL_CHECK(s && observed, WDL_ERR_ARG, "wdl_solver_set_observed: NULL argument");
  std::vector<float> h((size_t)s->P * s->N);
  for (size_t e = 0; e < h.size(); ++e) h[e] = (float)observed[e];
  WDL_HIP(hipSetDevice(s->ctx->device));
  WDL_HIP(hipMemcpyAsync(s->obs.p, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice,
                         s->ctx->stream));
  WDL_HIP(hipStreamSynchronize(s->ctx->stream));
  s->observed = true;
  return WDL_OK;
}

int wdl_solver_set_observed_device(wdl_solver* s, const float* d_observed) {
  WDL_CHECK(s && d_observed, WDL_ERR_ARG, "wdl_solver_set_observed_device: NULL argument");
  WDL_HIP(hipSetDevice(s->ctx->device));
  WDL_HIP(hipMemcpyAsync(s->obs.p, d_observed, (size_t)s->P * s->N * sizeof(float),
                         hipMemcpyDeviceToDevice, s->ctx->stream));
  s->observed = true;
  return WDL_OK;
}

int wdl_solver_set_problem(wdl_solver* s, const double* atoms, const double* lambdas) {
  WDL_CHECK(s && atoms && lambdas, WDL_ERR_ARG, "wdl_solver_set_problem: NULL argument");
  const int K = s->K, P = s->P;
  const long long N = s->N;
  // Problem::normalize_values (inverseWasserstein.h:53-74), per solve.
  s->hostLam.assign(lambdas, lambdas + (size_t)P * K);
  for (int p = 0; p < P; ++p) {
    double t = 0;
    for (int i = 0; i < K; ++i) t += s->hostLam[(size_t)p * K + i];
    for (int i = 0; i < K; ++i) s->hostLam[(size_t)p * K + i] /= t;
  }
  s->hostAtoms.assign(atoms, atoms + (size_t)K * N);
  for (int i = 0; i < K; ++i) {
    double t = 0;
    for (long long x = 0; x < N; ++x) t += s->hostAtoms[(size_t)i * N + x];
    for (long long x = 0; x < N; ++x) s->hostAtoms[(size_t)i * N + x] /= t;
  }
  std::vector<float> ha(s->hostAtoms.begin(), s->hostAtoms.end());
  std::vector<float> hl(s->hostLam.begin(), s->hostLam.end());
  WDL_HIP(hipSetDevice(s->ctx->device));
  if (s->flags & WDL_FLAG_LOG_DOMAIN) {
    // log(pdfs) of :704, :756, :796, from the FP64 atoms
    std::vector<float> hla(ha.size());
    for (size_t e = 0; e < hla.size(); ++e) hla[e] = (float)std::log(s->hostAtoms[e]);
    WDL_HIP(hipMemcpyAsync(s->logA.p, hla.data(), hla.size() * sizeof(float),
                           hipMemcpyHostToDevice, s->ctx->stream));
    WDL_HIP(hipStreamSynchronize(s->ctx->stream));
  }
  WDL_HIP(hipMemcpyAsync(s->A.p, ha.data(), ha.size() * sizeof(float), hipMemcpyHostToDevice,
                         s->ctx->stream));
  WDL_HIP(hipMemcpyAsync(s->lam.p, hl.data(), hl.size() * sizeof(float), hipMemcpyHostToDevice,
                         s->ctx->stream));
  WDL_HIP(hipStreamSynchronize(s->ctx->stream));
  s->problem = true;
  return WDL_OK;
}

int wdl_solver_set_problem_device(wdl_solver* s, const float* d_atoms, const float* d_lambdas) {
  WDL_CHECK(s && d_atoms && d_lambdas, WDL_ERR_ARG, "wdl_solver_set_problem_device: NULL argument");
  WDL_HIP(hipSetDevice(s->ctx->device));
  WDL_HIP(hipMemcpyAsync(s->A.p, d_atoms, (size_t)s->K * s->N * sizeof(float),
                         hipMemcpyDeviceToDevice, s->ctx->stream));
  WDL_HIP(hipMemcpyAsync(s->lam.p, d_lambdas, (size_t)s->P * s->K * sizeof(float),
                         hipMemcpyDeviceToDevice, s->ctx->stream));
  if (s->flags & WDL_FLAG_LOG_DOMAIN) {
    const long long KN = (long long)s->K * s->N;
    hipLaunchKernelGGL(k_log, dim3((unsigned)std::min<long long>(1024, (KN + 255) / 256)), dim3(256),
                       0, s->ctx->stream, s->A.as<float>(), s->logA.as<float>(), KN);
    WDL_HIP(hipGetLastError());
  }
  s->problem = true;
  return WDL_OK;
}

int wdl_solver_set_warm_state(wdl_solver* s, const double* b_storage) {
  WDL_CHECK(s && b_storage, WDL_ERR_ARG, "wdl_solver_set_warm_state: NULL argument");
  WDL_CHECK(s->flags & WDL_FLAG_WARM_RESTART, WDL_ERR_STATE,
            "wdl_solver_set_warm_state: solver created without WDL_FLAG_WARM_RESTART");
  const size_t n = (size_t)s->P * s->K * s->N;
  WDL_HIP(hipSetDevice(s->ctx->device));
  if (s->flags & WDL_FLAG_LOG_DOMAIN) {  // log b, FP64
    WDL_HIP(hipMemcpyAsync(s->bStore.p, b_storage, n * sizeof(double), hipMemcpyHostToDevice,
                           s->ctx->stream));
    WDL_HIP(hipStreamSynchronize(s->ctx->stream));
    return WDL_OK;
  }
  std::vector<float> h(n);
  for (size_t e = 0; e < n; ++e) h[e] = (float)b_storage[e];
  WDL_HIP(hipMemcpyAsync(s->bStore.p, h.data(), n * sizeof(float), hipMemcpyHostToDevice,
                         s->ctx->stream));
  WDL_HIP(hipStreamSynchronize(s->ctx->stream));
  return WDL_OK;
}

int wdl_solver_get_warm_state(wdl_solver* s, double* b_temp) {
  WDL_CHECK(s && b_temp, WDL_ERR_ARG, "wdl_solver_get_warm_state: NULL argument");
  WDL_CHECK(s->flags & WDL_FLAG_WARM_RESTART, WDL_ERR_STATE,
            "wdl_solver_get_warm_state: solver created without WDL_FLAG_WARM_RESTART");
  const size_t n = (size_t)s->P * s->K * s->N;
  WDL_HIP(hipSetDevice(s->ctx->device));
  if (s->flags & WDL_FLAG_LOG_DOMAIN) {  // log b, FP64
    WDL_HIP(hipMemcpyAsync(b_temp, s->bTemp.p, n * sizeof(double), hipMemcpyDeviceToHost,
                           s->ctx->stream));
    WDL_HIP(hipStreamSynchronize(s->ctx->stream));
    return WDL_OK;
  }
  std::vector<float> h(n);
  WDL_HIP(hipMemcpyAsync(h.data(), s->bTemp.p, n * sizeof(float), hipMemcpyDeviceToHost,
                         s->ctx->stream));
  WDL_HIP(hipStreamSynchronize(s->ctx->stream));
  for (size_t e = 0; e < n; ++e) b_temp[e] = h[e];
  return WDL_OK;
}

int wdl_solver_warm_commit(wdl_solver* s) {
  WDL_CHECK(s, WDL_ERR_ARG, "wdl_solver_warm_commit: NULL argument");
  WDL_CHECK(s->flags & WDL_FLAG_WARM_RESTART, WDL_ERR_STATE,
            "wdl_solver_warm_commit: solver created without WDL_FLAG_WARM_RESTART");
  WDL_HIP(hipSetDevice(s->ctx->device));
  const size_t esz = (s->flags & WDL_FLAG_LOG_DOMAIN) ? sizeof(double) : sizeof(float);
  WDL_HIP(hipMemcpyAsync(s->bStore.p, s->bTemp.p, (size_t)s->P * s->K * s->N * esz,
                         hipMemcpyDeviceToDevice, s->ctx->stream));
  return WDL_OK;
}

int wdl_solver_run(wdl_solver* s, int mode) {
  WDL_CHECK(s, WDL_ERR_ARG, "wdl_solver_run: NULL argument");
  WDL_CHECK(s->observed && s->problem, WDL_ERR_STATE,
            "wdl_solver_run: set_observed and set_problem must be called first");
  WDL_CHECK(mode == WDL_RUN_BATCHED || mode == WDL_RUN_PLAIN, WDL_ERR_ARG,
            "wdl_solver_run: unknown mode");
  WDL_HIP(hipSetDevice(s->ctx->device));
  int st;
  if (mode == WDL_RUN_PLAIN) {
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(1), 0, s->ctx->stream, s->events.as<unsigned>());
    for (int p = 0; p < s->P; ++p)
      if ((st = s->enqueue(p, 1, p == s->P - 1, s->ctx->stream))) return st;
  } else if (s->flags & WDL_FLAG_NO_GRAPH) {
    if ((st = s->enqueue_batched())) return st;
  } else {
    if (!s->captured) {
      // Two graphs, the first L/4 forward steps and the rest: a graph's first
      // kernel starts only after the runtime has prepared much of the graph
      // (at L = 100 with two histogram groups, ~0.6 ms of the ~1.2 ms the
      // launch call takes), so a small first graph starts at once and the
      // second is prepared while it runs.  A run waited for at once (an
      // L-BFGS evaluation, a projection round) gains that time.
      const int split = s->L >= 8 ? s->L / 4 : 0;
      const int imax = std::numeric_limits<int>::max();
      for (int k = split ? 0 : 1; k < 2; ++k) {
        s->op_lo = k == 0 ? 0 : split;
        s->op_hi = k == 0 ? split : imax;
        WDL_HIP(hipStreamBeginCapture(s->ctx->stream, hipStreamCaptureModeThreadLocal));
        st = s->enqueue_batched();
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(s->ctx->stream, &g);
        s->op_lo = 0;
        s->op_hi = imax;
        if (st || e != hipSuccess) {
          if (g) (void)hipGraphDestroy(g);
          s->release_graphs();
          if (st) return st;
          set_error(std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
          return WDL_ERR_HIP;
        }
        s->graph[k] = g;
        const hipError_t ie = hipGraphInstantiate(&s->exec[k], g, nullptr, nullptr, 0);
        if (ie != hipSuccess) {
          s->release_graphs();
          set_error(std::string("hipGraphInstantiate: ") + hipGetErrorString(ie));
          return WDL_ERR_HIP;
        }
      }
      s->captured = true;
    }
    for (int k = 0; k < 2; ++k)
      if (s->exec[k]) WDL_HIP(hipGraphLaunch(s->exec[k], s->ctx->stream));
  }
  s->ran = true;
  return WDL_OK;
}

int wdl_solver_run_stage(wdl_solver* s, int stage, int step) {
  WDL_CHECK(s, WDL_ERR_ARG, "wdl_solver_run_stage: NULL argument");
  WDL_CHECK(s->ran, WDL_ERR_STATE, "wdl_solver_run_stage: run the solver once first");
  WDL_CHECK(step >= 1 && step <= s->L, WDL_ERR_ARG, "wdl_solver_run_stage: step out of 1..L");
  WDL_CHECK(stage == WDL_STAGE_FORWARD || stage == WDL_STAGE_BACKWARD, WDL_ERR_ARG,
            "wdl_solver_run_stage: unknown stage");
  WDL_HIP(hipSetDevice(s->ctx->device));
  const SolveArgs a = s->args(0, s->P);
  const int lds = s->lds_bytes(a);  // all P solves, on the context stream
  const int gs = stage_grid(s->P, s->K);
  if (s->exact)
    return stage == WDL_STAGE_FORWARD ? exact_forward(a, s->ctx->stream, step)
                                      : exact_backward(a, s->ctx->stream, step);
  if (s->tiled)
    return stage == WDL_STAGE_FORWARD ? tiled_forward(a, s->ctx->stream, step)
                                      : tiled_backward(a, s->ctx->stream, step);
  if (stage == WDL_STAGE_FORWARD)
    launch_forward(a, gs, lds, s->ctx->stream, step);
  else
    launch_backward(a, gs, lds, s->ctx->stream, step);
  WDL_HIP(hipGetLastError());
  return WDL_OK;
}

int wdl_solver_get_results(wdl_solver* s, double* loss, double* grad_lambda, double* grad_dic_sum,
                           double* barycenters) {
  WDL_CHECK(s, WDL_ERR_ARG, "wdl_solver_get_results: NULL argument");
  WDL_CHECK(s->ran, WDL_ERR_STATE, "wdl_solver_get_results: nothing has run");
  WDL_HIP(hipSetDevice(s->ctx->device));
  hipStream_t st = s->ctx->stream;
  const int P = s->P, K = s->K;
  const long long N = s->N;
  std::vector<double> lp;
  std::vector<float> q;
  if (loss) {
    lp.resize((size_t)P * kLossBlocks);
    WDL_HIP(hipMemcpyAsync(lp.data(), s->lossPart.p, lp.size() * 8, hipMemcpyDeviceToHost, st));
  }
  if (grad_lambda)
    WDL_HIP(hipMemcpyAsync(grad_lambda, s->GL.p, (size_t)P * K * 8, hipMemcpyDeviceToHost, st));
  if (grad_dic_sum && !(s->flags & WDL_FLAG_WEIGHTS_ONLY))
    WDL_HIP(hipMemcpyAsync(grad_dic_sum, s->GDsum.p, (size_t)K * N * 8, hipMemcpyDeviceToHost, st));
  if (grad_dic_sum && (s->flags & WDL_FLAG_WEIGHTS_ONLY))
    std::memset(grad_dic_sum, 0, (size_t)K * N * sizeof(double));
  if (barycenters) {
    q.resize((size_t)P * N);
    WDL_HIP(hipMemcpyAsync(q.data(), s->Q.p, q.size() * 4, hipMemcpyDeviceToHost, st));
  }
  WDL_HIP(hipStreamSynchronize(st));
  if (loss)
    for (int p = 0; p < P; ++p) {
      double t = 0;
      for (int b = 0; b < kLossBlocks; ++b) t += lp[(size_t)p * kLossBlocks + b];
      loss[p] = t;
    }
  if (barycenters)
    for (size_t e = 0; e < q.size(); ++e) barycenters[e] = q[e];
  return WDL_OK;
}

int wdl_solver_log_range_events(wdl_solver* s, long long* count) {
  WDL_CHECK(s && count, WDL_ERR_ARG, "wdl_solver_log_range_events: NULL argument");
  unsigned h = 0;
  WDL_HIP(hipSetDevice(s->ctx->device));
  WDL_HIP(hipMemcpyAsync(&h, s->events.p, sizeof(unsigned), hipMemcpyDeviceToHost, s->ctx->stream));
  WDL_HIP(hipStreamSynchronize(s->ctx->stream));
  *count = h;
  return WDL_OK;
}

int wdl_solver_set_iterations(wdl_solver* s, const int* iterations) {
  WDL_CHECK(s, WDL_ERR_ARG, "wdl_solver_set_iterations: NULL argument");
  std::vector<int> lp((size_t)s->P, s->L);
  if (iterations)
    for (int p = 0; p < s->P; ++p) {
      WDL_CHECK(iterations[p] >= 0 && iterations[p] <= s->L, WDL_ERR_ARG,
                "wdl_solver_set_iterations: iterations must be in 0..L");
      lp[p] = iterations[p];
    }
  WDL_HIP(hipSetDevice(s->ctx->device));
  const int mx = *std::max_element(lp.begin(), lp.end());
  WDL_HIP(hipMemcpyAsync(s->Lp.p, lp.data(), lp.size() * sizeof(int), hipMemcpyHostToDevice,
                         s->ctx->stream));
  WDL_HIP(hipMemcpyAsync(s->lmax.p, &mx, sizeof(int), hipMemcpyHostToDevice, s->ctx->stream));
  WDL_HIP(hipStreamSynchronize(s->ctx->stream));
  s->hostLp = lp;
  return WDL_OK;
}

int wdl_solver_get_grad_dic(wdl_solver* s, double* grad_dic) {
  WDL_CHECK(s && grad_dic, WDL_ERR_ARG, "wdl_solver_get_grad_dic: NULL argument");
  WDL_CHECK(!(s->flags & WDL_FLAG_WEIGHTS_ONLY), WDL_ERR_STATE,
            "wdl_solver_get_grad_dic: the solver computes weight gradients only");
  WDL_CHECK(s->ran, WDL_ERR_STATE, "wdl_solver_get_grad_dic: nothing has run");
  const size_t n = (size_t)s->P * s->K * s->N;
  std::vector<float> h(n);
  WDL_HIP(hipSetDevice(s->ctx->device));
  WDL_HIP(hipMemcpyAsync(h.data(), s->GD.p, n * 4, hipMemcpyDeviceToHost, s->ctx->stream));
  WDL_HIP(hipStreamSynchronize(s->ctx->stream));
  for (size_t e = 0; e < n; ++e) grad_dic[e] = h[e];
  return WDL_OK;
}

int wdl_solver_set_tuning(wdl_solver* s, int nstreams, int fwd_threads, int bwd_threads,
                          int line_waves, int line_group) {
  WDL_CHECK(s, WDL_ERR_ARG, "wdl_solver_set_tuning: NULL argument");
  WDL_CHECK(nstreams >= 0 && nstreams <= 16, WDL_ERR_ARG,
            "wdl_solver_set_tuning: nstreams must be in 0..16");
  WDL_CHECK(fwd_threads == 0 || fwd_threads == 768 || fwd_threads == 1024, WDL_ERR_ARG,
            "wdl_solver_set_tuning: fwd_threads must be 0, 768 or 1024");
  WDL_CHECK(bwd_threads == 0 || bwd_threads == 512 || bwd_threads == 768, WDL_ERR_ARG,
            "wdl_solver_set_tuning: bwd_threads must be 0, 512 or 768");
  WDL_CHECK(line_waves == 0 || line_waves == 1 || line_waves == 2 || line_waves == 4, WDL_ERR_ARG,
            "wdl_solver_set_tuning: line_waves must be 0, 1, 2 or 4");
  WDL_CHECK(line_group == 0 || line_group == 4 || line_group == 8, WDL_ERR_ARG,
            "wdl_solver_set_tuning: line_group must be 0, 4 or 8");
  WDL_HIP(hipSetDevice(s->ctx->device));
  int st;
  if (nstreams && (st = s->set_groups(nstreams))) return st;
  const int f = fwd_threads ? fwd_threads : s->fwdT, b = bwd_threads ? bwd_threads : s->bwdT;
  const int len = std::max(s->W, s->H);
  const int w = line_waves ? xlse_waves_per_line(len, line_waves) : s->xwpl;
  const int g = line_group ? xlse_lines(len, xlse_waves_per_line(len, w), line_group) : s->xlw;
  if (f != s->fwdT || b != s->bwdT || w != s->xwpl || g != s->xlw) {
    if ((st = s->drop_graph())) return st;
    s->fwdT = f;
    s->bwdT = b;
    s->xwpl = w;
    s->xlw = g;
  }
  return WDL_OK;
}

int wdl_solver_get_tuning(wdl_solver* s, int* nstreams, int* fwd_threads, int* bwd_threads,
                          int* line_waves, int* line_group) {
  WDL_CHECK(s, WDL_ERR_ARG, "wdl_solver_get_tuning: NULL argument");
  if (nstreams) *nstreams = s->ngroups;
  const bool fp32 = !(s->flags & (WDL_FLAG_LOG_DOMAIN | WDL_FLAG_BF16X3)) && !s->tiled;
  if (fwd_threads) *fwd_threads = fp32 ? s->fwdT : 0;
  if (bwd_threads) *bwd_threads = fp32 ? s->bwdT : 0;
  const int len = std::max(s->W, s->H);
  if (line_waves) *line_waves = s->exact ? xlse_waves_per_line(len, s->xwpl) : 0;
  if (line_group)
    *line_group = s->exact ? xlse_lines(len, xlse_waves_per_line(len, s->xwpl), s->xlw) : 0;
  return WDL_OK;
}

// Times `reps` batched runs of the solver as it stands (after one run that
// captures the graph and warms up); -1 on error.
static float time_runs(wdl_solver* s, int reps, hipEvent_t e0, hipEvent_t e1) {
  if (wdl_solver_run(s, WDL_RUN_BATCHED)) return -1.f;
  hipStream_t st = s->ctx->stream;
  if (hipEventRecord(e0, st) != hipSuccess) return -1.f;
  for (int r = 0; r < reps; ++r)
    if (wdl_solver_run(s, WDL_RUN_BATCHED)) return -1.f;
  float ms = 0.f;
  if (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
      hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
    return -1.f;
  return ms;
}

int wdl_solver_autotune(wdl_solver* s, int reps, int* nstreams) {
  WDL_CHECK(s, WDL_ERR_ARG, "wdl_solver_autotune: NULL argument");
  WDL_CHECK(reps >= 1, WDL_ERR_ARG, "wdl_solver_autotune: reps must be >= 1");
  WDL_CHECK(s->observed && s->problem, WDL_ERR_STATE,
            "wdl_solver_autotune: set_observed and set_problem must be called first");
  WDL_HIP(hipSetDevice(s->ctx->device));
  hipEvent_t e0, e1;
  WDL_HIP(hipEventCreate(&e0));
  WDL_HIP(hipEventCreate(&e1));
  int rc = WDL_OK;
  // Coordinate descent over the launch shape: the histogram-group count
  // first, then the shape the solver form has (workgroup widths of the FP32
  // resident stages, or waves per line of the exact-shift passes), each
  // candidate timed over `reps` batched runs.
  struct Shape { int groups, fwd, bwd, wpl, lw; };
  // the shape in use, waves per line resolved (0 would keep a trial's value)
  const int len = std::max(s->W, s->H);
  auto current = [&]() {
    const int w = xlse_waves_per_line(len, s->xwpl);
    return Shape{s->ngroups, s->fwdT, s->bwdT, s->exact ? w : 0,
                 s->exact ? xlse_lines(len, w, s->xlw) : 0};
  };
  Shape best = current();
  float bestMs = -1.f;
  auto trial = [&](const Shape& c) {
    if (rc) return;
    if ((rc = wdl_solver_set_tuning(s, c.groups, c.fwd, c.bwd, c.wpl, c.lw))) return;
    const float ms = time_runs(s, reps, e0, e1);
    if (ms < 0.f) {
      set_error("wdl_solver_autotune: a timed run failed");
      rc = WDL_ERR_HIP;
      return;
    }
    if (bestMs < 0.f || ms < bestMs) {
      bestMs = ms;
      best = current();
    }
  };
  for (int n = 1; n <= 8 && n <= s->P; n *= 2) trial(Shape{n, best.fwd, best.bwd, best.wpl, best.lw});
  const bool fp32 = !(s->flags & (WDL_FLAG_LOG_DOMAIN | WDL_FLAG_BF16X3)) && !s->tiled;
  if (fp32) {
    const int shapes[][2] = {{1024, 768}, {768, 768}, {1024, 512}, {768, 512}};
    const Shape base = best;
    for (const auto& sh : shapes)
      if (sh[0] != base.fwd || sh[1] != base.bwd)
        trial(Shape{base.groups, sh[0], sh[1], base.wpl, base.lw});
  }
  if (s->exact) {
    const Shape base = best;
    for (int w : {1, 2, 4})
      if (xlse_waves_per_line(len, w) == w && w != base.wpl)
        trial(Shape{base.groups, base.fwd, base.bwd, w, base.lw});
    const Shape b2 = best;  // then 4 or 8 lines per workgroup at the best waves per line
    for (int g : {4, 8})
      if (xlse_lines(len, b2.wpl, g) == g && g != b2.lw)
        trial(Shape{b2.groups, b2.fwd, b2.bwd, b2.wpl, g});
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  if ((rc = wdl_solver_set_tuning(s, best.groups, best.fwd, best.bwd, best.wpl, best.lw))) return rc;
  if (nstreams) *nstreams = best.groups;
  return WDL_OK;
}

int wdl_solver_set_slot(wdl_solver* s, int slot, const double* observed, const double* lambdas,
                        int iterations, double step) {
  WDL_CHECK(s, WDL_ERR_ARG, "wdl_solver_set_slot: NULL argument");
  WDL_CHECK(slot >= 0 && slot < s->P, WDL_ERR_ARG, "wdl_solver_set_slot: slot out of 0..P-1");
  WDL_CHECK(iterations >= 0 && iterations <= s->L, WDL_ERR_ARG,
            "wdl_solver_set_slot: iterations must be in 0..L");
  WDL_CHECK(std::isfinite(step) && step >= 0, WDL_ERR_ARG,
            "wdl_solver_set_slot: step must be finite and >= 0");
  const int K = s->K;
  const long long N = s->N;
  hipStream_t st = s->ctx->stream;
  WDL_HIP(hipSetDevice(s->ctx->device));
  WDL_HIP(hipStreamSynchronize(st));  // the slot's arrays may be in use by a running round
  std::vector<float> ho, hl;
  if (observed) {
    ho.resize((size_t)N);
    for (long long x = 0; x < N; ++x) ho[(size_t)x] = (float)observed[x];
    WDL_HIP(hipMemcpyAsync(s->obs.as<float>() + (long long)slot * N, ho.data(), N * sizeof(float),
                           hipMemcpyHostToDevice, st));
    s->observed = true;
  }
  if (lambdas) {  // normalize_values (:53-74)
    double t = 0;
    for (int i = 0; i < K; ++i) t += lambdas[i];
    WDL_CHECK(t > 0 && std::isfinite(t), WDL_ERR_ARG, "wdl_solver_set_slot: weights must sum > 0");
    hl.resize((size_t)K);
    for (int i = 0; i < K; ++i) hl[(size_t)i] = (float)(lambdas[i] / t);
    WDL_HIP(hipMemcpyAsync(s->lam.as<float>() + (long long)slot * K, hl.data(), K * sizeof(float),
                           hipMemcpyHostToDevice, st));
  }
  const float et = (float)step;
  WDL_HIP(hipMemcpyAsync(s->Lp.as<int>() + slot, &iterations, sizeof(int), hipMemcpyHostToDevice, st));
  s->hostLp[(size_t)slot] = iterations;
  const int mx = *std::max_element(s->hostLp.begin(), s->hostLp.end());
  WDL_HIP(hipMemcpyAsync(s->lmax.p, &mx, sizeof(int), hipMemcpyHostToDevice, st));
  WDL_HIP(hipMemcpyAsync(s->eta.as<float>() + slot, &et, sizeof(float), hipMemcpyHostToDevice, st));
  if (s->flags & WDL_FLAG_WARM_RESTART) {
    // the start value of b_storage, 1 in both builds of the reference (:380-384)
    if (s->flags & WDL_FLAG_LOG_DOMAIN)
      hipLaunchKernelGGL(k_fill<double>, dim3(256), dim3(256), 0, st,
                         s->bStore.as<double>() + (long long)slot * K * N, 1.0, (long long)K * N);
    else
      hipLaunchKernelGGL(k_fill<float>, dim3(256), dim3(256), 0, st,
                         s->bStore.as<float>() + (long long)slot * K * N, 1.f, (long long)K * N);
    WDL_HIP(hipGetLastError());
  }
  WDL_HIP(hipStreamSynchronize(st));
  return WDL_OK;
}

int wdl_solver_projection_round(wdl_solver* s) {
  WDL_CHECK(s, WDL_ERR_ARG, "wdl_solver_projection_round: NULL argument");
  const int need = WDL_FLAG_WEIGHTS_ONLY | WDL_FLAG_WARM_RESTART;
  WDL_CHECK((s->flags & need) == need, WDL_ERR_STATE,
            "wdl_solver_projection_round: needs WDL_FLAG_WEIGHTS_ONLY | WDL_FLAG_WARM_RESTART");
  int st = wdl_solver_run(s, WDL_RUN_BATCHED);
  if (st) return st;
  hipStream_t str = s->ctx->stream;
  hipLaunchKernelGGL(k_project_update, dim3((s->P + 63) / 64), dim3(64), 0, str, s->lam.as<float>(),
                     s->GL.as<double>(), s->Lp.as<int>(), s->eta.as<float>(), s->P, s->K);
  const long long KN = (long long)s->K * s->N;
  const dim3 cg((unsigned)std::min<long long>(64, (KN + 255) / 256), (unsigned)s->P);
  if (s->flags & WDL_FLAG_LOG_DOMAIN)
    hipLaunchKernelGGL(k_warm_commit_active<double>, cg, dim3(256), 0, str, s->bStore.as<double>(),
                       s->bTemp.as<double>(), s->Lp.as<int>(), KN);
  else
    hipLaunchKernelGGL(k_warm_commit_active<float>, cg, dim3(256), 0, str, s->bStore.as<float>(),
                       s->bTemp.as<float>(), s->Lp.as<int>(), KN);
  WDL_HIP(hipGetLastError());
  return WDL_OK;
}

int wdl_solver_get_slot(wdl_solver* s, int slot, double* lambdas, double* loss) {
  WDL_CHECK(s, WDL_ERR_ARG, "wdl_solver_get_slot: NULL argument");
  WDL_CHECK(slot >= 0 && slot < s->P, WDL_ERR_ARG, "wdl_solver_get_slot: slot out of 0..P-1");
  const int K = s->K;
  std::vector<float> hl((size_t)K);
  std::vector<double> lp(kLossBlocks);
  hipStream_t st = s->ctx->stream;
  WDL_HIP(hipSetDevice(s->ctx->device));
  if (lambdas)
    WDL_HIP(hipMemcpyAsync(hl.data(), s->lam.as<float>() + (long long)slot * K, K * sizeof(float),
                           hipMemcpyDeviceToHost, st));
  if (loss)
    WDL_HIP(hipMemcpyAsync(lp.data(), s->lossPart.as<double>() + (long long)slot * kLossBlocks,
                           kLossBlocks * sizeof(double), hipMemcpyDeviceToHost, st));
  WDL_HIP(hipStreamSynchronize(st));
  if (lambdas)
    for (int i = 0; i < K; ++i) lambdas[i] = hl[(size_t)i];
  if (loss) {
    double t = 0;
    for (int b = 0; b < kLossBlocks; ++b) t += lp[(size_t)b];
    *loss = t;
  }
  return WDL_OK;
}

int wdl_gradient_both(wdl_solver* s, const double* atoms, const double* lambdas, double* loss,
                      double* grad_lambda, double* grad_dic_sum, double* barycenters) {
  int st = wdl_solver_set_problem(s, atoms, lambdas);
  if (st) return st;
  if ((st = wdl_solver_run(s, WDL_RUN_BATCHED))) return st;
  return wdl_solver_get_results(s, loss, grad_lambda, grad_dic_sum, barycenters);
}

int wdl_solver_set_scale(wdl_solver* s, double scale_dictionary, int exp_weight) {
  WDL_CHECK(s, WDL_ERR_ARG, "wdl_solver_set_scale: NULL argument");
  WDL_CHECK(scale_dictionary != 0 && std::isfinite(scale_dictionary), WDL_ERR_ARG,
            "wdl_solver_set_scale: scale_dictionary must be finite and non-zero");
  s->scaleDic = scale_dictionary;
  s->expWeight = exp_weight ? 1 : 0;
  return WDL_OK;
}

// evaluate_both with x and g in device memory: prologue, batched solve and
// epilogue on the context stream; *f (host) after a synchronisation.
static int evaluate_device(wdl_solver* s, const double* dx, double* dg, double* f) {
  WDL_CHECK(s->observed, WDL_ERR_STATE, "evaluate_both: set_observed must be called first");
  const int P = s->P, K = s->K;
  const long long N = s->N;
  hipStream_t st = s->ctx->stream;
  WDL_HIP(hipSetDevice(s->ctx->device));
  const bool logd = s->flags & WDL_FLAG_LOG_DOMAIN;
  WDL_CHECK(evaluate_prologue(dx, s->evW.as<double>(), s->evA.as<double>(), s->lam.as<float>(),
                              s->A.as<float>(), logd ? s->logA.as<float>() : nullptr,
                              s->evP.as<double>(), P, K, N, s->scaleDic, s->expWeight, st) == 0,
            WDL_ERR_HIP, "evaluate_both: prologue launch failed");
  s->problem = true;
  int rc = wdl_solver_run(s, WDL_RUN_BATCHED);
  if (rc) return rc;
  WDL_CHECK(evaluate_epilogue(s->GL.as<double>(), s->GDsum.as<double>(), s->lossPart.as<double>(),
                              s->evW.as<double>(), s->evA.as<double>(), s->evP.as<double>(), dg,
                              s->evF.as<double>(), P, K, N, s->scaleDic, s->expWeight,
                              (s->flags & WDL_FLAG_WEIGHTS_ONLY) ? 1 : 0, st) == 0,
            WDL_ERR_HIP, "evaluate_both: epilogue launch failed");
  WDL_HIP(hipMemcpyAsync(f, s->evF.p, sizeof(double), hipMemcpyDeviceToHost, st));
  WDL_HIP(hipStreamSynchronize(st));
  return WDL_OK;
}

int wdl_evaluate_both_device(wdl_solver* s, const double* d_x, double* d_g, double* f) {
  WDL_CHECK(s && d_x && d_g && f, WDL_ERR_ARG, "wdl_evaluate_both_device: NULL argument");
  return evaluate_device(s, d_x, d_g, f);
}

double wdl_evaluate_both(void* instance, const double* x, double* g, int n, double step) {
  (void)step;
  const double nan = std::numeric_limits<double>::quiet_NaN();
  wdl_solver* s = static_cast<wdl_solver*>(instance);
  if (!s || !x || !g) {
    set_error("wdl_evaluate_both: NULL argument");
    return nan;
  }
  const long long nn = (long long)s->P * s->K + (long long)s->K * s->N;
  if ((long long)n != nn) {
    set_error("wdl_evaluate_both: n != P*K + K*N");
    return nan;
  }
  // Only x and g cross PCIe; the reparametrisation (:998-1031), the batched
  // solve and the loss sum, projections and chain rule (:1089-1136) run on
  // the device.
  hipStream_t st = s->ctx->stream;
  double f = nan;
  if (hipSetDevice(s->ctx->device) != hipSuccess ||
      hipMemcpyAsync(s->evX.p, x, (size_t)nn * sizeof(double), hipMemcpyHostToDevice, st) !=
          hipSuccess) {
    set_error("wdl_evaluate_both: copy of x failed");
    return nan;
  }
  if (evaluate_device(s, s->evX.as<double>(), s->evG.as<double>(), &f)) return nan;
  if (hipMemcpyAsync(g, s->evG.p, (size_t)nn * sizeof(double), hipMemcpyDeviceToHost, st) !=
          hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess) {
    set_error("wdl_evaluate_both: copy of g failed");
    return nan;
  }
  return f;
}

}  // extern "C"

// ---- groups: uneven grid sizes in one launch per stage kernel ----------------
// The members' own buffers hold their problems and results; the group keeps
// their arguments and the block tables in device memory and replays one
// captured sequence: forward launches s = 1..Lmax, the final stage, backward
// launches t = Lmax..1 (start-aligned in both directions, as within one solver,
// so a member with fewer iterations sits out the launches past its own L),
// then, for projection rounds, the update and the warm commit.
struct wdl_group {
  wdl_ctx* ctx = nullptr;
  std::vector<wdl_solver*> ms;
  DevBuf dev;  // GroupMember[nm], then int starts[GK_COUNT][nm + 1]
  GroupArgs ga{};
  GroupShape sh{};
  bool projection = false;  // every member weights-only with a warm-restart store
  // [kind][part]: kind 0 run, 1 projection round; part 0 the first Lmax/4
  // forward launches (none when Lmax < 8), part 1 the rest -- a small first
  // graph starts at once and the second is prepared while it runs.
  hipGraph_t graph[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
  hipGraphExec_t exec[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
  bool captured[2] = {false, false};

  // The launch sequence numbered (forward steps, the final stage, backward
  // steps, the reductions / updates); only numbers in [lo, hi) are launched.
  int enqueue(bool proj, int lo, int hi) {
    hipStream_t st = ctx->stream;
    int e, op = 0;
    auto take = [&]() {
      const int o = op++;
      return o >= lo && o < hi;
    };
    for (int s = 1; s <= sh.Lmax; ++s)
      if (take() && (e = exact_group_forward(ga, sh, st, s))) return e;
    if (take())
      hipLaunchKernelGGL(k_final_log_group, dim3(sh.grid[GK_FINAL]), dim3(kThreads), 0, st, ga);
    for (int t = sh.Lmax; t >= 1; --t)
      if (take() && (e = exact_group_backward(ga, sh, st, t))) return e;
    if (!take()) return WDL_OK;
    if (!sh.weightsOnly)
      hipLaunchKernelGGL(k_reduce_gd_group, dim3(sh.grid[GK_RGD]), dim3(256), 0, st, ga);
    if (proj) {
      hipLaunchKernelGGL(k_project_update_group, dim3(sh.grid[GK_PROJ]), dim3(64), 0, st, ga);
      hipLaunchKernelGGL(k_warm_commit_group, dim3(sh.grid[GK_WARM]), dim3(256), 0, st, ga);
    }
    WDL_HIP(hipGetLastError());
    return WDL_OK;
  }

  void release_graphs(int k) {
    for (int h = 0; h < 2; ++h) {
      if (exec[k][h]) (void)hipGraphExecDestroy(exec[k][h]);
      if (graph[k][h]) (void)hipGraphDestroy(graph[k][h]);
      exec[k][h] = nullptr;
      graph[k][h] = nullptr;
    }
    captured[k] = false;
  }

  int launch(bool proj) {
    const int k = proj ? 1 : 0;
    if (!captured[k]) {
      const int split = sh.Lmax >= 8 ? sh.Lmax / 4 : 0;
      for (int h = split ? 0 : 1; h < 2; ++h) {
        WDL_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        const int st = enqueue(proj, h == 0 ? 0 : split,
                               h == 0 ? split : std::numeric_limits<int>::max());
        hipGraph_t gr = nullptr;
        const hipError_t e = hipStreamEndCapture(ctx->stream, &gr);
        if (st || e != hipSuccess) {
          if (gr) (void)hipGraphDestroy(gr);
          release_graphs(k);
          if (st) return st;
          set_error(std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
          return WDL_ERR_HIP;
        }
        graph[k][h] = gr;
        const hipError_t ie = hipGraphInstantiate(&exec[k][h], gr, nullptr, nullptr, 0);
        if (ie != hipSuccess) {
          release_graphs(k);
          set_error(std::string("hipGraphInstantiate: ") + hipGetErrorString(ie));
          return WDL_ERR_HIP;
        }
      }
      captured[k] = true;
    }
    for (int h = 0; h < 2; ++h)
      if (exec[k][h]) WDL_HIP(hipGraphLaunch(exec[k][h], ctx->stream));
    for (wdl_solver* m : ms) m->ran = true;
    return WDL_OK;
  }
};

extern "C" {

int wdl_group_create(wdl_ctx* ctx, wdl_solver* const* solvers, int n, wdl_group** out) {
  WDL_CHECK(ctx && solvers && out, WDL_ERR_ARG, "wdl_group_create: NULL argument");
  WDL_CHECK(n >= 1 && n <= 4096, WDL_ERR_ARG, "wdl_group_create: n must be in 1..4096");
  int wo = -1, maxlen = 0, wpl = 1, Lmax = 0;
  bool proj = true;
  for (int m = 0; m < n; ++m) {
    const wdl_solver* s = solvers[m];
    WDL_CHECK(s, WDL_ERR_ARG, "wdl_group_create: NULL solver");
    WDL_CHECK(s->ctx == ctx, WDL_ERR_ARG,
              "wdl_group_create: every member must be created on the group's context");
    WDL_CHECK(s->exact, WDL_ERR_UNSUPPORTED,
              "wdl_group_create: members must be exact-shift log-domain solvers "
              "(WDL_FLAG_LOG_DOMAIN | WDL_FLAG_EXACT_LSE)");
    WDL_CHECK(!s->ring, WDL_ERR_UNSUPPORTED,
              "wdl_group_create: members with checkpointed histories are not supported");
    WDL_CHECK(s->loss != WDL_LOSS_WASSERSTEIN, WDL_ERR_UNSUPPORTED,
              "wdl_group_create: the Wasserstein loss runs its own inner solve per solver");
    const int w = (s->flags & WDL_FLAG_WEIGHTS_ONLY) ? 1 : 0;
    WDL_CHECK(wo < 0 || wo == w, WDL_ERR_ARG,
              "wdl_group_create: members must all have WDL_FLAG_WEIGHTS_ONLY or none");
    wo = w;
    for (int k = 0; k < m; ++k)
      WDL_CHECK(solvers[k] != s, WDL_ERR_ARG, "wdl_group_create: a solver is listed twice");
    const int len = std::max(s->W, s->H);
    maxlen = std::max(maxlen, len);
    wpl = std::max(wpl, xlse_waves_per_line(len));
    Lmax = std::max(Lmax, s->L);
    proj = proj && w && (s->flags & WDL_FLAG_WARM_RESTART);
  }
  WDL_HIP(hipSetDevice(ctx->device));
  const int lds = xlse_lds_bytes(maxlen, kXWaves);
  int st;
  if ((st = xlse_group_set_attr(lds))) return st;
  std::vector<SolveArgs> as;
  std::vector<GroupMember> gm((size_t)n);
  for (int m = 0; m < n; ++m) {
    wdl_solver* s = solvers[m];
    as.push_back(s->args(0, s->P));
    gm[(size_t)m].a = as.back();
    gm[(size_t)m].ig = 1.0 / s->gamma;
    gm[(size_t)m].eta = s->eta.as<float>();
    gm[(size_t)m].GDsum = s->GDsum.as<double>();
    gm[(size_t)m].lmax = s->lmax.as<int>();
  }
  std::vector<int> counts[GK_COUNT];
  group_blocks(as, lds, ctx->num_cus, counts);
  std::vector<int> starts((size_t)GK_COUNT * (n + 1));
  GroupShape sh{};
  for (int k = 0; k < GK_COUNT; ++k) {
    long long acc = 0;
    for (int m = 0; m < n; ++m) {
      starts[(size_t)k * (n + 1) + m] = (int)acc;
      acc += counts[k][(size_t)m];
    }
    WDL_CHECK(acc < (1LL << 31), WDL_ERR_ARG, "wdl_group_create: too many workgroups");
    starts[(size_t)k * (n + 1) + n] = (int)acc;
    sh.grid[k] = (int)acc;
  }
  sh.wpl = wpl;
  sh.lds = lds;
  sh.maxlen = maxlen;
  sh.weightsOnly = wo;
  sh.Lmax = Lmax;
  wdl_group* g = new wdl_group();
  g->ctx = ctx;
  g->ms.assign(solvers, solvers + n);
  g->sh = sh;
  g->projection = proj;
  const size_t mbytes = (size_t)n * sizeof(GroupMember);
  if ((st = g->dev.ensure(mbytes + starts.size() * sizeof(int)))) {
    delete g;
    return st;
  }
  hipError_t e = hipMemcpy(g->dev.p, gm.data(), mbytes, hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy((char*)g->dev.p + mbytes, starts.data(), starts.size() * sizeof(int),
                  hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    g->dev.release();
    delete g;
    set_error(std::string("wdl_group_create: ") + hipGetErrorString(e));
    return WDL_ERR_HIP;
  }
  g->ga.m = g->dev.as<GroupMember>();
  g->ga.starts = reinterpret_cast<const int*>((char*)g->dev.p + mbytes);
  g->ga.nm = n;
  *out = g;
  return WDL_OK;
}

int wdl_group_destroy(wdl_group* g) {
  if (!g) return WDL_OK;
  (void)hipSetDevice(g->ctx->device);
  (void)hipStreamSynchronize(g->ctx->stream);
  for (int k = 0; k < 2; ++k) g->release_graphs(k);
  g->dev.release();
  delete g;
  return WDL_OK;
}

int wdl_group_run(wdl_group* g) {
  WDL_CHECK(g, WDL_ERR_ARG, "wdl_group_run: NULL argument");
  for (wdl_solver* s : g->ms)
    WDL_CHECK(s->observed && s->problem, WDL_ERR_STATE,
              "wdl_group_run: set_observed and set_problem must be called on every member");
  WDL_HIP(hipSetDevice(g->ctx->device));
  return g->launch(false);
}

int wdl_group_projection_round(wdl_group* g) {
  WDL_CHECK(g, WDL_ERR_ARG, "wdl_group_projection_round: NULL argument");
  WDL_CHECK(g->projection, WDL_ERR_STATE,
            "wdl_group_projection_round: every member needs WDL_FLAG_WEIGHTS_ONLY | "
            "WDL_FLAG_WARM_RESTART");
  for (wdl_solver* s : g->ms)
    WDL_CHECK(s->observed && s->problem, WDL_ERR_STATE,
              "wdl_group_projection_round: set_observed and set_problem must be called on "
              "every member");
  WDL_HIP(hipSetDevice(g->ctx->device));
  return g->launch(true);
}

int wdl_group_shape(const wdl_group* g, int* grids, int* waves_per_line, int* lds_bytes) {
  WDL_CHECK(g, WDL_ERR_ARG, "wdl_group_shape: NULL argument");
  if (grids) std::copy(g->sh.grid, g->sh.grid + GK_COUNT, grids);
  if (waves_per_line) *waves_per_line = g->sh.wpl;
  if (lds_bytes) *lds_bytes = g->sh.lds;
  return WDL_OK;
}

}  // extern "C"
