// Device-side pieces shared by the solver's translation units: the argument
// block of every stage kernel, workgroup -> (histogram, atom) mappings, the
// losses (loss.h) and a workgroup reduction.
#pragma once

#include "wdl_internal.h"
#include "wdl_logconv.h"

namespace wdl {

constexpr int kLossBlocks = 32;
constexpr int kWassersteinIters = 50;  // WassersteinLoss's n_iter default (loss.h:95)

struct SolveArgs {
  ConvGeom G;
  int K, P, L, p0, np, loss;
  long long N;      // pixels per image
  long long stage;  // P*K*N floats between consecutive history entries
  long long qstage; // P*N floats between consecutive Qh entries
  const float* tab;
  const float* A;
  const float* lam;
  const float* obs;
  float* Bh;
  float* Qh;  // standard domain: q^s per histogram, L x P x N
  float* Ch;
  float* Fh;
  float* V0;
  float* V1;
  float* R0;
  float* R1;
  float* GD;
  double* GL;
  float* Q;
  float* G0;
  double* lossPart;
  const float* bInit;   // warm-restart store b_storage / b_temp (standard domain, FP32)
  float* bSave;
  const double* lbInit;  // log domain: log b in FP64, as the histories
  double* lbSave;
  const int* Lp;      // per-histogram Sinkhorn iterations (<= L), or null for L
  // Checkpointed histories: `ring` history slots hold entry j at j % ring (0:
  // every entry has its own slot); forward launches skip histograms with fewer
  // than `rmin` iterations (a segment's recomputation runs only the histograms
  // whose backward reaches past it).
  int ring;
  int rmin;
  float* LF0;  // standard domain: safe_log phi^s of the previous forward stage
  float* LF1;  // (ping-pong; the V buffers themselves without checkpoints)
  int weightsOnly;    // skip the dictionary-gradient chain
  int fwdT, bwdT;     // threads per workgroup of the FP32 resident forward / backward stages
  int xwpl;           // waves per line of the exact-shift passes (0: by line length)
  int xlw;            // lines per workgroup of the exact-shift passes (4 or 8)
  int bf3;            // split-BF16x3 convolution passes (standard domain)
  const float* logA;  // K x N log atoms (log-domain solver)
  unsigned* events;   // lines whose spread exceeded kLogRange (log-domain solver)
  // Tiled solver (a side above kMaxSide): geometry, the inputs of the second
  // backward convolution (FP32; log domain: FP64 log magnitudes and sign
  // bytes), and per-tile partial weight-gradient dots [P][K][tiles].
  TileGeom T;
  float* X0;
  float* X1;
  unsigned char* XS0;
  unsigned char* XS1;
  double* GLt;
  double* gsumPart;  // [P][kLossBlocks] partial sums of g_L = loss'(q) q (k_final)
  double* Mt;        // tiled: [L][P][K][tiles] partial sums of log phi (forward)
  // Wasserstein loss: log a, log b of its inner Sinkhorn, a convolution result
  // and the transposed intermediate, P x N FP64 each
  double* wla;
  double* wlb;
  double* wlc;
  double* wlt;
  // exact-shift log domain: gamma, FP64 scratch images and sign bytes
  double gamma;
  double* xa;
  double* xt;
  double* xdp;  // per-chunk partial sums of the weight-gradient dots
  // Start of the forward's scratch images in xa / xt: the backward's images of
  // histogram p lie at p * chains * K * N, the forward's at xfwd + p * K * N
  // (a region of its own when the backward has two chains, so the forward of
  // one histogram group never writes what another group's backward reads).
  long long xfwd;
  // tiled log domain: the backward's cross-atom sums -sum_j v_j, -sum_j r_j (P x N)
  float* nj;
  float* gj;
  unsigned char* sga;
  unsigned char* sgt;
};

// Workgroup -> (histogram, atom).  With 8+ histograms the K workgroups of a
// histogram get block ids congruent mod 8, i.e. the same XCD.
__device__ inline bool map_block(const SolveArgs& a, int& p, int& i) {
  const int b = blockIdx.x;
  int pl;
  if (a.np >= 8) {
    const int xcd = b & 7, slot = b >> 3;
    pl = (slot / a.K) * 8 + xcd;
    i = slot % a.K;
  } else {
    pl = b / a.K;
    i = b % a.K;
  }
  if (pl >= a.np) return false;
  p = a.p0 + pl;
  return true;
}

// (A: SolveArgs, or a struct holding the fields these read.)
template <class A>
__device__ inline int steps_of(const A& a, int p) { return a.Lp ? a.Lp[p] : a.L; }

// Forward launch s runs histogram p when p has an iteration s and reaches rmin.
__device__ inline bool fwd_runs(const SolveArgs& a, int p, int s) {
  const int lp = steps_of(a, p);
  return s <= lp && lp >= a.rmin;
}

// Launch t of the backward (t = L..1) is step t of every histogram that has
// that many iterations: a histogram with L_p < L starts its backward at launch
// L_p, so histogram p's history entry of step s is written and read at the
// same launch index s in both directions.  Returns 0 when p has no step t.
template <class A>
__device__ inline int bwd_step(const A& a, int p, int t) {
  return t <= steps_of(a, p) ? t : 0;
}

// History slot of entry j (0-based: entry s-1 belongs to iteration s).
template <class A>
__host__ __device__ inline long long hslot(const A& a, int j) {
  return a.ring ? (long long)(j % a.ring) : (long long)j;
}

static int stage_grid(int np, int K) { return np >= 8 ? 8 * K * ((np + 7) / 8) : np * K; }

// Backward: two workgroups (dictionary chain, weight chain) per (histogram,
// atom), on block ids b and b + 8 (np >= 8) -- the same XCD, dispatched
// together -- so the arrays both chains read come from one L2.
__device__ inline bool map_block_bwd(const SolveArgs& a, int& p, int& i, int& chain) {
  const int b = blockIdx.x;
  int pl;
  if (a.weightsOnly) {  // one workgroup per (histogram, atom), the weight chain
    chain = 1;
    return map_block(a, p, i);
  }
  if (a.np >= 8) {
    const int xcd = b & 7, slot = b >> 3;
    chain = slot & 1;
    const int s2 = slot >> 1;
    pl = (s2 / a.K) * 8 + xcd;
    i = s2 % a.K;
  } else {
    chain = b & 1;
    pl = (b >> 1) / a.K;
    i = (b >> 1) % a.K;
  }
  if (pl >= a.np) return false;
  p = a.p0 + pl;
  return true;
}

__device__ inline float loss_grad(int loss, float q, float o) {
  if (loss == WDL_LOSS_KL) return safe_logf(q / o);          // loss.h:83-89
  if (loss == WDL_LOSS_TV) return (q - o > 0.f) ? 0.5f : -0.5f;  // loss.h:58-64
  return 2.f * (q - o);                                       // loss.h:34-40
}

__device__ inline double loss_term(int loss, float q, float o) {
  const double dq = q, dob = o;
  if (loss == WDL_LOSS_KL) {                                  // loss.h:74-81
    const double r = dq / dob;
    return dq * std::log(r > 1e-250 ? r : 1e-250) - dq + dob;
  }
  if (loss == WDL_LOSS_TV) return 0.5 * std::fabs(dq - dob);  // loss.h:49-56
  return (dq - dob) * (dq - dob);                             // loss.h:25-32
}

// The weight-gradient dot of backward step s (:649, :786), <log phi_i^s, g_s>,
// is evaluated as <log phi_i^s - m, g_s> + m E_s with m the image mean of
// log phi_i^s and E_s = sum_x g_s taken exact: summing the recursion of
// :655-672 over the pixels gives sum_x r_i^{s-1} = sum_x r_i^s - lam_i sum_x g_s,
// so with sum_i lam_i = 1, E_{L-1} = -sum_x g_L and E_s = 0 for s <= L-2.  Same
// value in exact arithmetic; in FP32 the rounding of g no longer enters
// multiplied by the mean of log phi (a large offset common to all atoms),
// which otherwise dominates the error of dL/dlambda.  Returns E_s for s < L_p.
__device__ inline double exact_gsum(const SolveArgs& a, int p, int s) {
  if (s <= steps_of(a, p) - 2) return 0.0;
  double t = 0;
  for (int b = 0; b < kLossBlocks; ++b) t += a.gsumPart[(long long)p * kLossBlocks + b];
  return -t;
}

__device__ inline double vsum_d(float x) { return (double)x; }
__device__ inline double vsum_d(floatx4 x) { return (double)x.x + x.y + x.z + x.w; }
__device__ inline double vsum_d(double x) { return x; }
__device__ inline double vsum_d(doublex4 x) { return x.x + x.y + x.z + x.w; }

// Tiled solver launchers (one Sinkhorn stage = several launches), in
// wdl_solver_tiled.hip.
int tiled_lds_bytes(const SolveArgs& a);
int tiled_set_lds_attr();
int tiled_forward(const SolveArgs& a, hipStream_t st, int s);
int tiled_backward(const SolveArgs& a, hipStream_t st, int t);
int tiled_reduce_gl(const SolveArgs& a, hipStream_t st);

// Exact-shift log-domain stages (wdl_xlse.hip).
int xlse_set_lds_attr(int maxlen);
int exact_forward(const SolveArgs& a, hipStream_t st, int s);
int exact_backward(const SolveArgs& a, hipStream_t st, int t);

// ---- groups: one launch per stage kernel over several solvers -----------------
// The members of a group are exact-shift log-domain solvers of any grid size,
// atom count, histogram count, iteration count and gamma.  Every stage kernel
// of a group launch is the concatenation of the members' grids: for each kind
// of launch, starts[kind][m] is member m's first workgroup (starts[kind][nm] the
// total), and a workgroup runs the single-solver kernel's body for its member
// with its block index inside that member's share.
struct GroupMember {
  SolveArgs a;          // the member's arguments over all its histograms (p0 = 0)
  double ig;            // 1 / gamma, as the single-solver launches compute it (host)
  const float* eta;     // mirror-descent steps per histogram (projection rounds)
  double* GDsum;        // dL/datoms summed over the histograms (K x N)
  const int* lmax;      // the largest per-histogram iteration count (device, kept by
                        // the solver): a member sits out the launches past it
};

enum GroupKind {
  GK_XF1 = 0,   // forward: log b^{s-1}, the first convolution's input
  GK_FA,        // forward passes along x (nimg = P K images of H lines)
  GK_FB,        // forward passes along y (P K images of W lines)
  GK_FINAL,     // k_final_log: kLossBlocks workgroups per histogram
  GK_DOT,       // weight-gradient dot chunks
  GK_XB1,       // both chains: the first backward convolution's input
  GK_DOTFIN,    // the dots' in-order finish
  GK_BA,        // backward passes along x (P nch K images)
  GK_BB,        // backward passes along y
  GK_PROJ,      // projection update (64 histograms per workgroup)
  GK_WARM,      // per-histogram warm commit
  GK_RGD,       // summed atom gradient
  GK_COUNT
};

struct GroupArgs {
  const GroupMember* m;  // device, nm entries
  const int* starts;     // device, GK_COUNT x (nm + 1)
  int nm;
};

// Member of workgroup b in a launch of `kind`, with its block index (lb) and
// block count (nb) inside the member's share.  starts[kind] is non-decreasing
// with starts[kind][nm] > b, so the search ends on a member with lb < nb.
__device__ inline int group_member(const GroupArgs& g, int kind, int b, int& lb, int& nb) {
  const int* st = g.starts + kind * (g.nm + 1);
  int lo = 0, hi = g.nm;  // st[lo] <= b < st[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (st[mid] <= b)
      lo = mid;
    else
      hi = mid;
  }
  lb = b - st[lo];
  nb = st[lo + 1] - st[lo];
  return lo;
}

// Host side of a group launch: the grid of every kind, the passes' waves per
// line (the most any member's lines need) and their LDS (the longest line's).
struct GroupShape {
  int grid[GK_COUNT];
  int wpl;
  int lds;
  int maxlen;
  int weightsOnly;
  int Lmax;
};

// Per-member block counts of every kind (host): what the single-solver
// launches use, with the passes' workgroups shared out over the members by
// their work when the group holds more than the device runs at once.
void group_blocks(const std::vector<SolveArgs>& ms, int lds, int cus,
                  std::vector<int> (&counts)[GK_COUNT]);
int xlse_group_set_attr(int lds);
int exact_group_forward(const GroupArgs& g, const GroupShape& sh, hipStream_t st, int s);
int exact_group_backward(const GroupArgs& g, const GroupShape& sh, hipStream_t st, int t);
// WassersteinLoss at the barycenters: fills G0, lossPart, gsumPart (wdl_xlse.hip).
int wasserstein_loss(const SolveArgs& a, hipStream_t st);
constexpr int kXMaxLen = 1024;  // longest line of the exact-shift passes
constexpr int kXWaves = 4;  // lines per workgroup by default (4 or 8; see xlse_lines)
constexpr int kXPad = 32;         // -inf positions on both sides of a staged line
constexpr int kXBPad = 16;        // -inf positions on both sides of its prefix / suffix maxima

// Dynamic LDS of k_xlse_pass: per line its values (FP64 scaled by log2 e, FP32),
// input signs, prefix and suffix maxima, outputs and output signs; per
// workgroup the penalty tables t^2 / gamma (FP32) and t^2 log2(e) / gamma (FP64)
// for t = -kXPad..kXPad.
constexpr int kXSym = 2 * kXPad + 1;
__host__ __device__ inline int xlse_lds_bytes(int len, int lw = kXWaves) {
  const int lp = len + 2 * kXPad, lb = len + 2 * kXBPad;
  return lw * (lp * 8 + len * 8 + lp * 4 * 2 + lb * 4 * 2 + len) + kXSym * 12 + 64;
}

// Lines per workgroup of a pass: 8 when asked for and the workgroup fits
// (at most 1024 threads, its LDS within a CU's), else 4.
constexpr int kXLdsMax = 160 * 1024;
inline int xlse_lines(int len, int wpl, int want) {
  return (want == 8 && 8 * 64 * wpl <= 1024 && xlse_lds_bytes(len, 8) <= kXLdsMax) ? 8 : 4;
}

// Waves per line of k_xlse_pass: long lines share their workgroup's LDS over
// more waves (the LDS per line is fixed by len), for latency hiding.  The
// default by length, or a requested count of 1, 2 or 4 (a lane stages at most
// 4 values of a line, so lines longer than 256 * wpl take more waves).
__host__ __device__ inline int xlse_waves_per_line(int len, int want = 0) {
  const int need = len <= 256 ? 1 : (len <= 512 ? 2 : 4);
  if (want == 1 || want == 2 || want == 4) return want < need ? need : want;
  return len <= 128 ? 1 : (len <= 512 ? 2 : 4);
}

// Pixels per workgroup of the exact-shift weight-gradient dot, and chunks per image.
constexpr int kXDotChunk = 4096;
__host__ __device__ inline long long xdot_chunks(long long N) {
  return (N + kXDotChunk - 1) / kXDotChunk;
}

// evaluate_both around the batched solve, on the device (wdl_evaluate.hip):
// x -> normalised weights / atoms (FP64 w64, a64 and the solve's FP32 copies),
// and the loss sum, projections and chain rule into g and *f (device).
// part: the per-chunk atom partial sums, 2 x K x ev_chunks(N) doubles.
constexpr int kEvChunk = 2048;
__host__ __device__ inline long long ev_chunks(long long N) {
  return (N + kEvChunk - 1) / kEvChunk;
}
int evaluate_prologue(const double* x, double* w64, double* a64, float* lam, float* A,
                      float* logA, double* part, int P, int K, long long N, double scale,
                      int expw, hipStream_t st);
int evaluate_epilogue(const double* GL, const double* GDsum, const double* lossPart,
                      const double* w64, const double* a64, double* part, double* g, double* f,
                      int P, int K, long long N, double scale, int expw, int weightsOnly,
                      hipStream_t st);

// Sum over the workgroup (NW waves), result valid in thread 0.
template <int NW = kWaves>
__device__ inline double block_sum(double v) {
  __shared__ double red[NW];
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double t = 0;
  if (threadIdx.x == 0)
    for (int w = 0; w < NW; ++w) t += red[w];
  __syncthreads();
  return t;
}

}  // namespace wdl
