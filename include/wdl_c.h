/*
 * wdl_c.h -- C ABI of libwdl, the gfx950 (MI355X) Sinkhorn-barycenter hot path
 * of Wasserstein dictionary learning.
 *
 * Every entry point returns an int status (WDL_OK = 0, < 0 on error; the text
 * is in wdl_last_error()), except wdl_evaluate_both, which has the
 * lbfgs_evaluate_t shape (lbfgs/lbfgs.h:378-384) and reports errors as NaN.
 * Signatures use plain pointers and sizes only.  Host-pointer entry points
 * take row-major H x W images, x fastest, nvectors of them back to back, in
 * FP64 -- the layout the reference passes to its KernelType (kernels.h:669).
 * Device-pointer entry points ("_device") take FP32 device memory and are
 * ordered on the context's stream (wdl_ctx_stream).
 *
 * Citations are into matthieuheitz/WassersteinDictionaryLearning, cpp/.
 */
#ifndef WDL_C_H
#define WDL_C_H

#ifdef __cplusplus
extern "C" {
#endif

#define WDL_OK 0
#define WDL_ERR_ARG (-1)
#define WDL_ERR_HIP (-2)
#define WDL_ERR_UNSUPPORTED (-3)
#define WDL_ERR_STATE (-4)
#define WDL_ERR_NOMEM (-5)

/* Loss indices of the reference CLI (main_dictionary_learning.cpp:189-197). */
#define WDL_LOSS_TV 1
#define WDL_LOSS_QUADRATIC 2
#define WDL_LOSS_WASSERSTEIN 3
#define WDL_LOSS_KL 4

/* Flags for wdl_gauss2d_create / wdl_solver_create. */
#define WDL_FLAG_LOG_DOMAIN 1   /* solver runs compute_gradient_both_log_signArray */
#define WDL_FLAG_WARM_RESTART 2 /* solver keeps b_storage / b_temp (inverseWasserstein.h:380-387) */
#define WDL_FLAG_DENSE 4        /* do not skip all-zero FP32 factor chunks */
#define WDL_FLAG_NO_GRAPH 8     /* batched runs launch eagerly instead of replaying a hipGraph */
#define WDL_FLAG_WEIGHTS_ONLY 16 /* only dL/dlambda (:635-673): projection onto a fixed dictionary */
#define WDL_FLAG_TILED 32       /* 64x64-tile kernels at any size (taken anyway above 128 per side) */
#define WDL_FLAG_EXACT_LSE 64   /* log domain: exact per-output maximum (kernels.h:1082-1091) for
                                   any spread of log values; FP32 sums leave out terms below e^-25 */
#define WDL_FLAG_BF16X3 128     /* standard domain, sides <= 128: the convolution's MFMA passes in
                                   split-BF16x3 (hi.hi + hi.lo + lo.hi) instead of FP32 */
#define WDL_FLAG_CHECKPOINT 256 /* solver: keep the b / conv_b / phi histories (inverseWasserstein.h:
                                   551-558) for a ring of ceil(sqrt(L)) + 2 iterations plus phi at
                                   every ceil(sqrt(L))-th iteration, and recompute each segment's
                                   forward during the backward (about one forward pass more);
                                   results are bit-identical to the full histories */

/* Run modes of wdl_solver_run. */
#define WDL_RUN_BATCHED 0 /* all P solves advance together, one launch per Sinkhorn stage */
#define WDL_RUN_PLAIN 1   /* one solve per launch sequence, P in turn (inverseWasserstein.h:1052) */

typedef struct wdl_ctx wdl_ctx;
typedef struct wdl_gauss2d wdl_gauss2d;
typedef struct wdl_gauss1d wdl_gauss1d;
typedef struct wdl_gauss3d wdl_gauss3d;
typedef struct wdl_solver wdl_solver;

const char* wdl_last_error(void);
const char* wdl_version(void);
/* Debug: while on (nonzero), every solver created fills the device buffers it
 * does not initialise itself (histories, scratch images, steps not yet set ...)
 * with 0xFF bytes (NaN in FP32 and FP64) before use, so a read of memory no
 * launch has written shows in the results.  No counterpart in the reference;
 * for tests.  Returns the previous setting. */
int wdl_debug_fill(int on);

/* ---- context: one device, one stream --------------------------------------- */
int wdl_ctx_create(int device, wdl_ctx** out);
int wdl_ctx_destroy(wdl_ctx* ctx);
/* The hipStream_t all work of this context is ordered on. */
void* wdl_ctx_stream(wdl_ctx* ctx);
int wdl_ctx_synchronize(wdl_ctx* ctx);

/* ---- KernelType: the separable Gaussian operator ---------------------------
 * Replaces Gaussian2DKernel (kernels.h:630-708), GaussianSSE2DKernel
 * (kernels.h:820-907) and LogSignArrayGaussian2DKernel (kernels.h:1034-1202)
 * as the object behind the reference's duck-typed KernelType.              */

/* Gaussian2DKernel(gamma, W, H): kernels.h:632-641. */
int wdl_gauss2d_create(wdl_ctx* ctx, int W, int H, double gamma, int flags, wdl_gauss2d** out);
int wdl_gauss2d_destroy(wdl_gauss2d* k);
/* kernel1d[d] = max(1e-250, exp(-d^2/gamma)), d < max(W,H): kernels.h:636-640. */
int wdl_gauss2d_table(const wdl_gauss2d* k, double* out);
/* Gaussian2DKernel::convolve / convolveAdjoint: kernels.h:669-702. */
int wdl_gauss2d_convolve(wdl_gauss2d* k, const double* u, double* out, int nvectors);
/* LogSignArrayGaussian2DKernel::log_convolve / log_convolveAdjoint: kernels.h:1073-1115. */
int wdl_gauss2d_log_convolve(wdl_gauss2d* k, const double* log_u, double* out, int nvectors);
/* LogSignArrayGaussian2DKernel::log_convolve_signArray / ...Adjoint: kernels.h:1117-1162.
 * signs: bit i of byte i/8 (LSB first), 1 <=> value >= 0 (signArray.h:8-22);
 * read as input signs, overwritten with output signs. */
int wdl_gauss2d_log_convolve_signarray(wdl_gauss2d* k, const double* log_u,
                                       unsigned char* signs, double* out, int nvectors);
/* FP32 device form of convolve (no equivalent in the reference; the batched path's operator). */
int wdl_gauss2d_convolve_device(wdl_gauss2d* k, const float* d_u, float* d_out, int nvectors);
/* Allocation-free form of wdl_gauss2d_log_convolve (d_signs == NULL) and
 * wdl_gauss2d_log_convolve_signarray (d_signs = the packed sign array, LSB
 * first, read and overwritten as there), on FP64 device memory, asynchronous
 * on the context stream.  No counterpart in the reference: it is for callers
 * that cannot allocate device memory (code under HIP graph capture, serving
 * loops, harnesses that forbid hipMalloc).  It calls no allocator: the
 * transposed intermediate and the sign bytes live in one static device array
 * per GPU (WDL_NOALLOC_MAX_VALUES * 11 bytes, 22 MiB, held while the library is
 * loaded), so nvectors*W*H may not exceed WDL_NOALLOC_MAX_VALUES (WDL_ERR_ARG
 * otherwise).  Every call on one GPU shares that array: calls that overlap on
 * one GPU (several streams or host threads at once) race; serialise them. */
#define WDL_NOALLOC_MAX_VALUES 2097152
int wdl_gauss2d_log_convolve_noalloc(wdl_gauss2d* k, const double* d_log_u, unsigned char* d_signs,
                                     double* d_out, int nvectors);

/* ---- 1-D and 3-D KernelTypes --------------------------------------------------
 * Replace Gaussian1DKernel (kernels.h:501-556) and Gaussian3DKernel
 * (kernels.h:712-815), the operators of 1-D and volumetric histograms.  Same
 * factor table, FP32 MFMA passes.  Lines longer than 128 values (W of the 1-D
 * kernel, D of the 3-D one) need a band of at most 114 nonzero FP32 factors
 * (gamma up to ~125); W, D <= 2048.  Larger W x H slices of the 3-D kernel take
 * the tiled 2-D passes. */
/* Gaussian1DKernel(gamma, W): kernels.h:503-511. */
int wdl_gauss1d_create(wdl_ctx* ctx, int W, double gamma, wdl_gauss1d** out);
int wdl_gauss1d_destroy(wdl_gauss1d* k);
/* Gaussian1DKernel::convolve / convolveAdjoint (nvectors vectors of W): kernels.h:533-552. */
int wdl_gauss1d_convolve(wdl_gauss1d* k, const double* u, double* out, int nvectors);
/* FP32 device form (no equivalent in the reference). */
int wdl_gauss1d_convolve_device(wdl_gauss1d* k, const float* d_u, float* d_out, int nvectors);
/* Gaussian3DKernel(gamma, W, H, D): kernels.h:714-724.  Voxel (x, y, z) at
 * z*W*H + y*W + x. */
int wdl_gauss3d_create(wdl_ctx* ctx, int W, int H, int D, double gamma, wdl_gauss3d** out);
int wdl_gauss3d_destroy(wdl_gauss3d* k);
/* Gaussian3DKernel::convolve / convolveAdjoint: kernels.h:762-810. */
int wdl_gauss3d_convolve(wdl_gauss3d* k, const double* u, double* out, int nvectors);
/* FP32 device form (no equivalent in the reference). */
int wdl_gauss3d_convolve_device(wdl_gauss3d* k, const float* d_u, float* d_out, int nvectors);

/* ---- Sinkhorn barycenter + unrolled gradient, P histograms per call --------
 * Holds the state WassersteinRegression keeps per solve (b, conv_b, phi
 * histories: inverseWasserstein.h:1214) for all P histograms at once.      */
int wdl_solver_create(wdl_ctx* ctx, int W, int H, int K, int P, int L, double gamma, int loss,
                      int flags, wdl_solver** out);
int wdl_solver_destroy(wdl_solver* s);
/* Device bytes wdl_solver_create would allocate (admission by footprint). */
long long wdl_solver_footprint(int W, int H, int K, int P, int L, int flags);
/* Problem::observed_pdf (P x N), inverseWasserstein.h:32-42. */
int wdl_solver_set_observed(wdl_solver* s, const double* observed);
int wdl_solver_set_observed_device(wdl_solver* s, const float* d_observed);
/* Atoms (K x N) and barycentric weights (P x K); normalised on the host as
 * Problem::normalize_values does per solve (inverseWasserstein.h:53-74). */
int wdl_solver_set_problem(wdl_solver* s, const double* atoms, const double* lambdas);
/* Already-normalised FP32 device atoms (K x N) and weights (P x K). */
int wdl_solver_set_problem_device(wdl_solver* s, const float* d_atoms, const float* d_lambdas);
/* Warm restart (flag WDL_FLAG_WARM_RESTART): b_storage (P x K x N) seeds b^0
 * of every solve (inverseWasserstein.h:1057-1061); every run saves b^L into
 * b_temp (:1081-1085); commit copies b_temp into b_storage (:418-422).  The
 * store starts at 1 (:380-384).  In the log domain it holds log b, as the
 * reference's log build does, kept in FP64 on the device (FP32 otherwise). */
int wdl_solver_set_warm_state(wdl_solver* s, const double* b_storage);
int wdl_solver_get_warm_state(wdl_solver* s, double* b_temp);
int wdl_solver_warm_commit(wdl_solver* s);
/* Batched runs split the P histograms into nstreams groups, each advancing
 * on its own stream (default min(4, P/16), at least 1).  No counterpart in
 * the reference; a tuning knob of the batched launch. */
int wdl_solver_set_streams(wdl_solver* s, int nstreams);
/* The launch shape of batched runs: histogram groups (streams, 1..16), threads
 * per workgroup of the standard-domain FP32 resident forward stage (768 or 1024)
 * and backward stage (512 or 768), waves per line of the exact-shift passes (1,
 * 2 or 4; raised to what the line length needs) and their lines per workgroup
 * (4 or 8; 4 where 8 would not fit).  0 keeps a value.  Results do not depend
 * on it.  No counterpart in the reference; set by the tuner below. */
int wdl_solver_set_tuning(wdl_solver* s, int nstreams, int fwd_threads, int bwd_threads,
                          int line_waves, int line_group);
/* The launch shape in use; a field the solver form does not use reads 0.  Any
 * pointer may be NULL. */
int wdl_solver_get_tuning(wdl_solver* s, int* nstreams, int* fwd_threads, int* bwd_threads,
                          int* line_waves, int* line_group);
/* Tunes the launch shape by coordinate descent, timing `reps` batched runs per
 * candidate: the histogram-group count (1, 2, 4, 8, at most P), then the
 * workgroup widths of the FP32 resident stages (forward x backward: 1024/768,
 * 768/768, 1024/512, 768/512) or the exact-shift passes' waves per line and
 * then lines per workgroup; keeps
 * the fastest and writes the group count to *nstreams (may be NULL; the rest
 * through wdl_solver_get_tuning).  Runs the solve, so results of earlier runs
 * are overwritten (the warm-restart store is not committed).  No counterpart in
 * the reference; the batched launch's tuner. */
int wdl_solver_autotune(wdl_solver* s, int reps, int* nstreams);
/* Per-histogram Sinkhorn iterations (0..L each; NULL = L for all): histogram p
 * runs iterations[p] forward steps and as many backward steps (the reference
 * has one Niters for all, inverseWasserstein.h:546; uneven counts serve
 * requests that arrive with their own budgets).  0 leaves histogram p idle:
 * no launch touches it, its loss reads 0 and its other outputs keep their
 * values (zero on a new solver), including its share of the summed atom
 * gradient. */
int wdl_solver_set_iterations(wdl_solver* s, const int* iterations);
/* Enqueue the forward L iterations and both unrolled backwards for all P
 * solves (compute_gradient_both, inverseWasserstein.h:544-674, or its log
 * form :678-814).  Asynchronous on the context stream. */
int wdl_solver_run(wdl_solver* s, int mode);
/* Launch one stage kernel of a batched run over all P solves: stage 0 is
 * forward iteration `step` (1..L), stage 1 is backward step `step` (L..1).
 * No counterpart in the reference; for timing and profiling a single kernel.
 * Stages read what earlier stages of a full run wrote. */
#define WDL_STAGE_FORWARD 0
#define WDL_STAGE_BACKWARD 1
int wdl_solver_run_stage(wdl_solver* s, int stage, int step);
/* Synchronise and fetch: loss per histogram (loss.h), dL/dlambda (P x K, the
 * resultLa of each solve), dL/datoms summed over histograms (K x N) and the
 * barycenters (P x N).  Any pointer may be NULL. */
int wdl_solver_get_results(wdl_solver* s, double* loss, double* grad_lambda,
                           double* grad_dic_sum, double* barycenters);
/* Range events of the last run.  Log domain (shifted MFMA form): image lines
 * whose spread of log values exceeded 80 nats, the range within which the
 * shifted FP32 MFMA pass loses no term that matters (the exact-shift form,
 * WDL_FLAG_EXACT_LSE, has no such limit and counts none).  Standard domain:
 * pixels whose final scalings phi or barycenter were not finite and positive,
 * i.e. the Sinkhorn scalings left the FP32 range (run such inputs in the log
 * domain).  0 means every value stayed in range.
 * No counterpart in the reference (FP64, exact per-output max). */
int wdl_solver_log_range_events(wdl_solver* s, long long* count);
/* resultDic of every solve (P x K x N). */
int wdl_solver_get_grad_dic(wdl_solver* s, double* grad_dic);
/* ---- continuous batching: per-histogram slots ------------------------------
 * No counterpart in the reference (its P solves are one serial loop,
 * inverseWasserstein.h:1052); these let a caller change one histogram of a
 * batched solver between runs, as a serving loop admits and retires
 * requests. */
/* Replaces histogram `slot`'s observed histogram (N values, NULL = keep) and
 * weights (K values, normalised here; NULL = keep), sets its Sinkhorn
 * iterations (0..L; 0 = idle: no launch touches it, its outputs keep their
 * values) and its mirror-descent step (wdl_solver_projection_round), and resets
 * its warm-restart store to the reference's start value 1 (:380-384).
 * Synchronous: waits for work on the context stream first. */
int wdl_solver_set_slot(wdl_solver* s, int slot, const double* observed, const double* lambdas,
                        int iterations, double step);
/* One projection round for every non-idle histogram (flags WDL_FLAG_WEIGHTS_ONLY |
 * WDL_FLAG_WARM_RESTART): the batched weights-only solve (:635-673) seeded from the
 * warm-restart store, then lambda_i <- lambda_i exp(-step g_i / max|g|) / Z with
 * g = resultLa - <lambda, resultLa> (the projection of :1095-1099), then the
 * per-histogram warm commit b_storage <- b_temp (:418-422).  A slot whose step
 * was never set (wdl_solver_set_slot) steps 0: its weights are only renormalised.
 * Asynchronous. */
int wdl_solver_projection_round(wdl_solver* s);
/* Histogram `slot`'s current weights (K) and the loss of its last run.  Synchronous. */
int wdl_solver_get_slot(wdl_solver* s, int slot, double* lambdas, double* loss);

/* ---- groups: solvers of uneven grid sizes in one launch per stage ----------
 * No counterpart in the reference (its solves are one serial loop over
 * histograms of one size, inverseWasserstein.h:1052-1113).  A group runs the
 * batched solves of several solvers -- each with its own W x H, K, P, L, gamma,
 * loss and per-histogram iteration counts -- as one launch per stage kernel:
 * every launch is the concatenation of the members' grids, and each
 * workgroup finds its member in a device-side block table.  Members must be
 * exact-shift log-domain solvers (WDL_FLAG_LOG_DOMAIN | WDL_FLAG_EXACT_LSE) on
 * the group's context, without WDL_FLAG_CHECKPOINT or the Wasserstein loss,
 * and either all or none WDL_FLAG_WEIGHTS_ONLY.  Launch s of the forward runs
 * the members (and histograms) with at least s iterations; the backward is
 * start-aligned the same way.  Results land in each member, exactly as
 * wdl_solver_run (WDL_RUN_BATCHED) of that member alone would leave them, and
 * are read with the member's own getters.  Destroy a group before its
 * members. */
typedef struct wdl_group wdl_group;
#define WDL_GROUP_LAUNCH_KINDS 12
int wdl_group_create(wdl_ctx* ctx, wdl_solver* const* solvers, int n, wdl_group** out);
int wdl_group_destroy(wdl_group* g);
/* compute_gradient_both_log_signArray (inverseWasserstein.h:678-814), or its
 * weight chain alone, for every histogram of every member.  Asynchronous on
 * the context stream (a captured graph, replayed on later calls). */
int wdl_group_run(wdl_group* g);
/* wdl_solver_projection_round for every member at once (members need
 * WDL_FLAG_WEIGHTS_ONLY | WDL_FLAG_WARM_RESTART).  Asynchronous. */
int wdl_group_projection_round(wdl_group* g);
/* The launch shape: workgroups of each of the WDL_GROUP_LAUNCH_KINDS launch
 * kinds (grids may be NULL), the passes' waves per line and their LDS bytes. */
int wdl_group_shape(const wdl_group* g, int* grids, int* waves_per_line, int* lds_bytes);

/* ---- served projection loop ----------------------------------------------
 * No counterpart in the reference.  A scheduler thread serves requests that
 * project histograms onto registered dictionaries (weights fitting by
 * projection rounds), batching the requests of one dictionary into pools of
 * slots that they join and leave between rounds (continuous batching),
 * running pools of different dictionaries concurrently on their own streams,
 * and admitting pools by device-memory footprint (wdl_solver_footprint)
 * within hbm_budget_bytes. */
typedef struct wdl_server wdl_server;
int wdl_server_create(int device, long long hbm_budget_bytes, int max_slots_per_pool,
                      wdl_server** out);
int wdl_server_destroy(wdl_server* sv);
/* How rounds are batched.  WDL_SERVER_POOLS (the default): each pool (one
 * dictionary and iteration class) runs its rounds on its own stream.
 * WDL_SERVER_MIXED: the pools of one iteration class share a stream and each
 * round runs all of them -- dictionaries of different grid sizes, atom counts
 * and gamma -- as one launch per stage kernel (wdl_group_projection_round);
 * requests join and leave a class's pools between its rounds.  Results are
 * the same in both modes.  Call before the first wdl_server_submit
 * (WDL_ERR_STATE afterwards). */
#define WDL_SERVER_POOLS 0
#define WDL_SERVER_MIXED 1
int wdl_server_set_batching(wdl_server* sv, int mode);
/* A dictionary: K atoms (K x W x H, normalised per atom by the solver), gamma,
 * loss, the largest iteration count a request may ask for, flags 0,
 * WDL_FLAG_LOG_DOMAIN or WDL_FLAG_LOG_DOMAIN | WDL_FLAG_EXACT_LSE.  Returns its id. */
int wdl_server_add_dictionary(wdl_server* sv, int W, int H, int K, double gamma, int loss,
                              int max_iterations, int flags, const double* atoms, int* dict_id);
/* Queue a request (the histogram is copied): `rounds` projection rounds of
 * `iterations` Sinkhorn iterations each, mirror-descent step `step`, weights
 * starting uniform.  Returns immediately with the request id. */
int wdl_server_submit(wdl_server* sv, int dict_id, const double* histogram, int iterations,
                      int rounds, double step, long long* request_id);
/* Wait up to timeout_s for a request; on completion writes its weights (K), the
 * loss of its last round, its latency (submit to completion) and its queueing
 * time (submit to admission), in ms, and forgets it.  Returns 1 on timeout. */
int wdl_server_wait(wdl_server* sv, long long request_id, double timeout_s, double* weights,
                    double* loss, double* latency_ms, double* queue_ms);
/* out6: completed requests, pools created, pools released, requests that waited
 * for device memory, device bytes held by pools, rounds launched. */
int wdl_server_stats(wdl_server* sv, long long* out6);
/* 0 when all work enqueued on the context has completed, 1 while some is pending. */
int wdl_ctx_query(wdl_ctx* ctx);

/* compute_gradient_both for every histogram: set_problem + run + get_results. */
int wdl_gradient_both(wdl_solver* s, const double* atoms, const double* lambdas, double* loss,
                      double* grad_lambda, double* grad_dic_sum, double* barycenters);

/* ---- the L-BFGS callback ------------------------------------------------------
 * WassersteinRegression::evaluate_both (inverseWasserstein.h:986-1145) with
 * the per-histogram loop (:1052-1113) batched on the GPU.  `instance` is a
 * wdl_solver*; x and g hold P*K weights then K*N atoms.  Pass it to lbfgs()
 * as the lbfgs_evaluate_t.  scale_dictionary and exp_weight play the roles
 * of WassersteinRegression::scaleDictionary and exp_weight. */
int wdl_solver_set_scale(wdl_solver* s, double scale_dictionary, int exp_weight);
/* The reparametrisation (:998-1031), the batched solve and the loss sum,
 * simplex-tangent projections and chain rule (:1089-1136) all run on the
 * device; x is copied in and g out (P*K + K*N doubles each), nothing else
 * crosses PCIe.  Returns f, NaN on error. */
double wdl_evaluate_both(void* instance, const double* x, double* g, int n, double step);
/* evaluate_both with x and g in device memory (FP64, P*K weights then K*N
 * atoms, ordered on the context stream); *f (host) receives the objective.
 * Synchronous.  No counterpart in the reference, whose x and g are host
 * arrays of lbfgs (lbfgs.h:378-384); for callers that keep the optimiser's
 * state on the device. */
int wdl_evaluate_both_device(wdl_solver* s, const double* d_x, double* d_g, double* f);

/* ---- histogram files (host only) ---------------------------------------------
 * ExportHistogramToCSV::exportHistogram (histogramIO.h:62-78): <folder>/<name>.csv,
 * one value per line in the stream's default 6-significant-digit form.
 * ExportHistogramToRAW::exportHistogram (histogramIO.h:208-220): <folder>/<name>.raw,
 * the n doubles in host byte order.  0 on success, -1 when the file cannot be
 * opened (message in wdl_last_error), WDL_ERR_ARG on null arguments. */
int wdl_export_histogram_csv(const char* folder, const char* name, const double* values,
                             long long n);
int wdl_export_histogram_raw(const char* folder, const char* name, const double* values,
                             long long n);
/* exportHistogramsBatch (histogramIO.h:80-97, 222-238): count histograms of
 * lengths[i] values; every file is attempted, -1 if any failed. */
int wdl_export_histograms_csv(const char* folder, const char* const* names,
                              const double* const* values, const long long* lengths, int count);
int wdl_export_histograms_raw(const char* folder, const char* const* names,
                              const double* const* values, const long long* lengths, int count);
/* load_csv_to_pdf (histogramIO.h:314-337): reads the file's values, divides them
 * by their sum and returns the sum, or -1 when the file cannot be opened.
 * *n_out receives the number of values read (the reference's loop counts a
 * trailing-whitespace read once more, with the last value); the first
 * min(*n_out, capacity) normalised values go to out (out may be NULL when
 * capacity is 0, to size a second call). */
double wdl_load_csv_to_pdf(const char* filename, double* out, long long capacity,
                           long long* n_out);

#ifdef __cplusplus
}
#endif
#endif /* WDL_C_H */
