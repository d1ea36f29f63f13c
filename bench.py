"""Benchmark: batched Sinkhorn barycenter + unrolled gradient on one MI355X.

Workload (one "step"): compute_gradient_both for all 64 histograms of the
batched standard-domain configuration -- 128x128 FP32 grid, K=5 atoms,
gamma=1, L=100 Sinkhorn iterations, quadratic loss -- the forward L
iterations and both unrolled backwards, with inputs already resident in HBM.
Metric: Sinkhorn-iterations/s = P * L * steps / seconds.

Also measured in the same run: the plain form (same kernels, one solve per
launch sequence, the reference's serial P loop), the dominant kernel's
launch duration for the roofline, and the reference CPU path on a bounded
sample (cpu_baseline), which also gives the error of the GPU result
against the reference's FP64 output.

python bench.py [--gpus 1] [--steps 20] [--warmup 3]
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import wassersteindictionarylearning_amd as wdl  # noqa: E402
from wassersteindictionarylearning_amd import _hip, synthetic  # noqa: E402

METRIC = ("Sinkhorn barycenter+grad iters/s (batched vs plain, 1 GPU) at ≤1e-5 err; "
          "p99 ms")
PEAK_FP32_MFMA = 157.3   # TFLOP/s, MI355X dense FP32 matrix
PEAK_HBM = 8000.0        # GB/s


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def band_halfwidth(table64) -> int:
    t32 = np.asarray(table64).astype(np.float32)
    nz = np.nonzero(t32)[0]
    return int(nz.max()) if nz.size else 0


def pass_flops(nrows, ncols, D):
    """MFMA FLOPs one 1-D pass executes: the 4-wide contraction steps holding a
    factor that can be nonzero (indices c0 - D .. c0 + 15 + D), for every 16 x 16
    output block (rows_gemm); one 16x16x4 MFMA per step and row block."""
    ncb = (ncols + 15) // 16
    nk4 = 4 * ncb
    rb = (nrows + 15) // 16
    steps = 0
    for cb in range(ncb):
        c0 = 16 * cb
        lo = max(0, c0 - D) // 4
        hi = min(nk4 - 1, (c0 + 15 + D) // 4)
        steps += hi - lo + 1
    return steps * rb * 16 * 16 * 4 * 2


def conv_flops(n, D):
    """(executed, dense 4n^3) FLOPs of one image-convolution."""
    return 2 * pass_flops(n, n, D), 4 * n ** 3


GOLDEN = os.path.join(ROOT, "tests", "golden")


def reference_fixture(name):
    """Outputs of the reference's FP64 log-domain solver on this benchmark's own
    inputs (tests/golden/large_*.npz, written by tests/golden/make_large.py from
    the compiled reference), or None when absent."""
    path = os.path.join(GOLDEN, name + ".npz")
    if not os.path.exists(path):
        return None
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def rel_err(a, b):
    a, b = np.asarray(a, np.float64).ravel(), np.asarray(b, np.float64).ravel()
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def projection_leg(ctx, requests=4096, chunk=512, one_launch=True):
    """Uneven batched projection: a fixed dictionary (K=8 blob atoms, 128x128),
    4096 histograms, per-request Sinkhorn iterations uniform in 20..100, weights
    gradient only, b seeded from the warm-restart store.  The scalings of this
    workload diverge (far transport at gamma=1 between sharp blobs: the standard
    domain overflows FP32 within ~10 iterations and FP64 within ~90), so it runs
    in the log domain with exact per-output shifts.  The first `requests` of
    the 4096 run, in chunks of `chunk` (the unrolled log histories of 512
    requests take 165 GB), each launched eagerly once.  The weight gradients of
    the requests the reference fixture holds are compared with it.  With
    one_launch, all requests also run as one batched solve with checkpointed
    histories (235 GB for 4096 requests instead of 1.3 TB in full)."""
    c = synthetic.make_problem("projection")
    n, K, L, g = c["n"], c["K"], c["L"], c["gamma"]
    P = min(requests, c["P"])
    rng = np.random.default_rng(7)
    atoms = synthetic.blob_histograms(rng, n, K, c["nblobs"], c["sigma"])  # the dictionary
    flags = (wdl.FLAG_WEIGHTS_ONLY | wdl.FLAG_WARM_RESTART | wdl.FLAG_LOG_DOMAIN |
             wdl.FLAG_EXACT_LSE | wdl.FLAG_NO_GRAPH)
    s = wdl.SinkhornSolver(ctx, n, n, K, chunk, L, g, c["loss"], flags)
    stream = ctx.stream
    e0, e1 = _hip.Event(), _hip.Event()
    total_ms, iters, finite = 0.0, 0, 0
    fx = reference_fixture("large_projection")
    errs = {}
    tuning = None
    for c0 in range(0, P, chunk):
        sl = slice(c0, c0 + chunk)
        s.set_observed(c["obs"][sl])
        s.set_problem(atoms, c["weights"][sl])
        s.set_iterations(c["iterations"][sl])
        if tuning is None:
            # the launch shape, tuned on the first chunk; 3 timed runs per
            # candidate (with one, 2 and 4 groups fell within its noise)
            s.autotune(3)
            tuning = s.tuning()
            log("projection tuning: %s" % tuning)
        ctx.synchronize()
        e0.record(stream)
        s.run()
        e1.record(stream)
        e1.synchronize()
        total_ms += e0.elapsed_ms(e1)
        iters += int(c["iterations"][sl].sum())
        r = s.results(barycenters=False)
        finite += int((np.isfinite(r["grad_lambda"]).all(1) & np.isfinite(r["loss"])).sum())
        if fx is not None:
            for j, p in enumerate(fx["requests"]):
                if c0 <= p < c0 + chunk:
                    errs[int(p)] = rel_err(r["grad_lambda"][p - c0], fx["grad_lambda"][j])
        log("projection chunk %d: %.1f ms" % (c0 // chunk, e0.elapsed_ms(e1)))
    s.close()
    one = None
    if one_launch:
        s = wdl.SinkhornSolver(ctx, n, n, K, P, L, g, c["loss"], flags | wdl.FLAG_CHECKPOINT)
        s.set_tuning(tuning["streams"], 0, 0, tuning["line_waves"], tuning["line_group"])
        s.set_observed(c["obs"][:P])
        s.set_problem(atoms, c["weights"][:P])
        s.set_iterations(c["iterations"][:P])
        ctx.synchronize()
        e0.record(stream)
        s.run()
        e1.record(stream)
        e1.synchronize()
        ms1 = e0.elapsed_ms(e1)
        r = s.results(barycenters=False)
        errs1 = {}
        if fx is not None:
            for j, p in enumerate(fx["requests"]):
                if p < P:
                    errs1[int(p)] = rel_err(r["grad_lambda"][p], fx["grad_lambda"][j])
        one = {"requests": P, "ms": round(ms1, 3),
               "value": round(int(c["iterations"][:P].sum()) / (ms1 / 1e3), 1),
               "unit": "Sinkhorn-iters/s",
               "hbm_footprint_gb": round(wdl.SinkhornSolver.footprint(
                   n, n, K, P, L, flags | wdl.FLAG_CHECKPOINT) / 1e9, 2),
               "finite_results": int((np.isfinite(r["grad_lambda"]).all(1)
                                      & np.isfinite(r["loss"])).sum()),
               "err_vs_fp64_grad_lambda": max(errs1.values()) if errs1 else None,
               "form": "all requests in one batched solve, histories checkpointed "
                       "(ring of ceil(sqrt(L))+2 iterations, segments recomputed in the backward)"}
        log("projection, one launch of %d (checkpointed): %.1f ms" % (P, ms1))
        s.close()
    e0.destroy()
    e1.destroy()
    return {"workload": "projection: 4096 requests, 128x128, K=8 fixed blob dictionary, gamma=1, "
                        "L per request uniform 20..100, weights gradient only, warm-restart seed, "
                        "log domain with exact per-output shifts",
            "requests": P, "of": c["P"], "chunk": chunk, "sinkhorn_iters": iters, "finite_results": finite,
            "ms": round(total_ms, 3), "value": round(iters / (total_ms / 1e3), 1),
            "unit": "Sinkhorn-iters/s", "tuning": tuning, "one_launch": one,
            "err_vs_fp64": {"grad_lambda": max(errs.values()) if errs else None,
                            "requests_checked": sorted(errs),
                            "against": "reference compute_gradient_both_log_signArray (FP64, "
                                       "tests/golden/large_projection.npz)"}}


def served_leg(duration_s=20.0, rate=2.0, rounds=3, step=0.5, budget=200 << 30, seed=11,
               mixed=False):
    """Served projection loop: Poisson arrivals at `rate` requests/s for
    `duration_s`, grid sizes 64/128/256/512 (probabilities .4/.3/.2/.1), one
    random-blob dictionary per size (K uniform 4..10, gamma from {0.5, 1, 2}),
    L uniform 20..500; each request fits its weights with `rounds` projection
    rounds (log domain, exact shifts: blob dictionaries make the scalings
    diverge).  The native server batches requests continuously within a
    device-memory budget; latency is submit to completion.  mixed: the pools
    of one iteration class (all grid sizes) run each round as one group
    launch instead of on their own streams."""
    flags = wdl.FLAG_LOG_DOMAIN | wdl.FLAG_EXACT_LSE
    sv = wdl.ProjectionServer(0, budget, 16, mixed=mixed)
    wl, arrivals, rng = synthetic.served_workload(seed, rate, duration_s)
    sizes = sorted(wl)
    dicts = {}
    for n in sizes:
        K, g, atoms = wl[n]
        dicts[n] = (sv.add_dictionary(atoms, n, n, g, wdl.LOSS_QUADRATIC, 500, flags), K, g)
    # one warm-up request per size (pool creation, graph capture) before the clock
    warm = [sv.submit(dicts[n][0], synthetic.blob_histograms(rng, n, 1, (1, 3), (3, 10))[0], 20, 1,
                      step) for n in sizes]
    for h in warm:
        sv.wait(h, 600.0)
    st0 = sv.stats()
    t0 = time.perf_counter()
    handles = []
    for (ta, n, L, h) in arrivals:
        while True:
            now = time.perf_counter() - t0
            if now >= ta:
                break
            time.sleep(min(ta - now, 0.002))
        handles.append(sv.submit(dicts[n][0], h, L, rounds, step))
    res = [sv.wait(hd, 900.0) for hd in handles]
    wall = time.perf_counter() - t0
    fx = reference_fixture("large_served_rounds")
    err = None
    if fx is not None and rounds == int(fx["rounds"]) and step == float(fx["step"]):
        ew = [rel_err(res[int(k)][0], fx["weights"][j]) for j, k in enumerate(fx["requests"])]
        el = [abs(res[int(k)][1] - fx["loss"][j]) / abs(fx["loss"][j])
              for j, k in enumerate(fx["requests"])]
        err = {"weights": max(ew), "loss": max(el),
               "weights_per_request": [float("%.3g" % e) for e in ew],
               "weights_bar": 5e-5,
               "weights_bar_why": "each round's exponentiated-gradient step divides the projected "
                                  "gradient by its largest component (up to ~13x below max "
                                  "|dL/dlambda| on these requests), so the gradients' error "
                                  "reaches the weights amplified; the losses stay within 1e-5",
               "requests_checked": fx["requests"].tolist(),
               "against": "the same projection rounds driven by the reference's FP64 "
                          "compute_gradient_both_log_signArray (tests/golden/large_served_rounds.npz)"}
    st = sv.stats()
    sv.close()
    lat = np.array([r[2] for r in res])
    qu = np.array([r[3] for r in res])
    iters = sum(a[2] for a in arrivals) * rounds
    finite = int(sum(np.isfinite(r[0]).all() and np.isfinite(r[1]) for r in res))
    by_size = {}
    for (a, r) in zip(arrivals, res):
        by_size.setdefault(a[1], []).append(r[2])
    return {"workload": "served projection: Poisson arrivals %.1f/s for %.0f s, sizes 64/128/256/512 "
                        "(p .4/.3/.2/.1), blob dictionaries K 4..10, gamma {0.5,1,2}, L 20..500, "
                        "%d rounds per request, log domain exact shifts" % (rate, duration_s, rounds),
            "batching": "mixed: one group launch per iteration class and round" if mixed else
                        "pools: one stream per dictionary and iteration class",
            "requests": len(res), "finite_results": finite, "wall_s": round(wall, 3),
            "solves_per_s": round(len(res) / wall, 3),
            "value": round(iters / wall, 1), "unit": "Sinkhorn-iters/s",
            "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 2),
                           "p99": round(float(np.percentile(lat, 99)), 2),
                           "by_size_p50": {str(k): round(float(np.percentile(v, 50)), 2)
                                           for k, v in sorted(by_size.items())}},
            "queue_ms": {"p50": round(float(np.percentile(qu, 50)), 2),
                         "p99": round(float(np.percentile(qu, 99)), 2)},
            "memory_waits": st["memory_waits"] - st0["memory_waits"],
            "pools_created": st["pools_created"], "pools_released": st["pools_released"],
            "rounds_launched": st["rounds"] - st0["rounds"], "hbm_budget_gb": budget / 1e9,
            "err_vs_fp64": err}


def uneven_leg(reps=3, seed=11):
    """Uneven grid sizes in one launch: projection rounds (weights only, warm
    start, exact shifts) over the served workload's four dictionaries at once
    -- 64x64, 128x128, 256x256, 512x512 with their own K and gamma -- with 32,
    16, 8 and 4 requests whose Sinkhorn iterations are uniform in 20..100.  A
    round is timed as one solver group (one launch per stage kernel over all
    sizes) and as the four solvers on streams of their own (the server's
    pooled form), same inputs; the results are compared bit for bit."""
    wl, _, _ = synthetic.served_workload(seed)
    counts = {64: 32, 128: 16, 256: 8, 512: 4}
    flags = (wdl.FLAG_WEIGHTS_ONLY | wdl.FLAG_WARM_RESTART | wdl.FLAG_LOG_DOMAIN |
             wdl.FLAG_EXACT_LSE)
    rng = np.random.default_rng(seed + 1)
    probs = {}
    for n, P in counts.items():
        K, g, atoms = wl[n]
        probs[n] = (K, g, atoms, synthetic.blob_histograms(rng, n, P, (1, 5), (3, 20)),
                    rng.integers(20, 101, P).astype(np.int32))
    iters = int(sum(int(v[4].sum()) for v in probs.values()))

    def build(ctxs):
        ss = []
        for (n, P), c in zip(counts.items(), ctxs):
            K, g, atoms, obs, it = probs[n]
            s = wdl.SinkhornSolver(c, n, n, K, P, 100, g, wdl.LOSS_QUADRATIC, flags)
            s.set_observed(obs)
            s.set_problem(atoms, np.full((P, K), 1.0 / K))
            for p in range(P):  # iterations and the served loop's mirror-descent step
                s.set_slot(p, None, None, int(it[p]), 0.5)
            s.set_streams(1)
            ss.append(s)
        return ss

    def timed(run, sync):
        run()  # warm-up: graph capture
        sync()
        t0 = time.perf_counter()
        for _ in range(reps):
            run()
        sync()
        return (time.perf_counter() - t0) / reps

    def state(ss):
        return [np.concatenate([s.get_slot(p)[0] for p in range(s.P)]) for s in ss]

    # one group on one context
    ctx = wdl.Context(0)
    ms = build([ctx] * 4)
    grp = wdl.SolverGroup(ctx, ms)
    tg = timed(grp.projection_round, ctx.synchronize)
    shape = grp.shape()
    wg = state(ms)
    grp.close()
    ctx.close()
    # the same solvers, each on its own context (stream)
    ctxs = [wdl.Context(0) for _ in counts]
    ss = build(ctxs)

    def run_all():
        for s in ss:
            s.projection_round()

    def sync_all():
        for c in ctxs:
            c.synchronize()

    tp = timed(run_all, sync_all)
    wp = state(ss)
    for c in ctxs:
        c.close()
    same = all(np.array_equal(a, b) for a, b in zip(wg, wp))
    return {"workload": "uneven grid sizes: projection rounds over 64^2 x32, 128^2 x16, 256^2 x8, "
                        "512^2 x4 requests (the served workload's dictionaries, K %s, gamma %s), "
                        "L_p uniform 20..100, weights only, warm start, exact shifts"
                        % ([probs[n][0] for n in counts], [probs[n][1] for n in counts]),
            "sinkhorn_iters_per_round": iters,
            "group": {"ms_per_round": round(tg * 1e3, 3), "value": round(iters / tg, 1),
                      "unit": "Sinkhorn-iters/s", "waves_per_line": shape["waves_per_line"],
                      "form": "one solver group: one launch per stage kernel over all sizes"},
            "streams": {"ms_per_round": round(tp * 1e3, 3), "value": round(iters / tp, 1),
                        "unit": "Sinkhorn-iters/s",
                        "form": "one solver per size, each on its own stream"},
            "results_bit_equal": bool(same)}


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_backward_batched_std_traffic.md")


def measured_traffic():
    """HBM bytes per k_backward launch (64 histograms) from the committed
    rocprofv3 --pmc summary (FETCH_SIZE and WRITE_SIZE, copy-calibrated)."""
    try:
        with open(TRAFFIC_FILE) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return int(line.split(":")[1]), os.path.relpath(TRAFFIC_FILE, ROOT)
    except OSError:
        pass
    return None, None


def log_leg(ctx, steps=2):
    """Log-domain small-epsilon workload: 256x256 (tiled kernels: the image
    exceeds one workgroup's LDS), 32 curve histograms, K=6, gamma=0.5, L=200,
    KL loss; a step is compute_gradient_both_log_signArray for all 32."""
    c = synthetic.make_problem("log_small_eps")
    n, P, K, L, g, loss = c["n"], c["P"], c["K"], c["L"], c["gamma"], c["loss"]
    s = wdl.SinkhornSolver(ctx, n, n, K, P, L, g, loss, wdl.FLAG_LOG_DOMAIN)
    s.set_observed(c["obs"])
    s.set_problem(c["atoms"], c["weights"])
    stream = ctx.stream
    s.autotune(1)
    tuning = s.tuning()
    s.run()  # warm-up and graph capture
    ctx.synchronize()
    e0, e1 = _hip.Event(), _hip.Event()
    e0.record(stream)
    for _ in range(steps):
        s.run()
    e1.record(stream)
    e1.synchronize()
    ms = e0.elapsed_ms(e1) / steps
    stage = {}
    for name, st in (("forward", wdl.STAGE_FORWARD), ("backward", wdl.STAGE_BACKWARD)):
        s.run_stage(st, L // 2)
        e0.record(stream)
        for _ in range(5):
            s.run_stage(st, L // 2)
        e1.record(stream)
        e1.synchronize()
        stage[name + "_stage_ms"] = round(e0.elapsed_ms(e1) / 5, 4)
    events = s.log_range_events()
    fx = reference_fixture("large_c3")
    err = None
    if fx is not None and int(fx["L"]) == L:
        s.run()  # the stage timings above overwrote results of the last full run
        r = s.results()
        gd = s.grad_dic()
        eb = [rel_err(r["bary"][0], fx["bary"])]
        el = [rel_err(r["grad_lambda"][0], fx["grad_lambda"])]
        hist = [0]
        fm = reference_fixture("large_c3_more")  # histograms 1-3: dL/dlambda, barycenter
        if fm is not None and int(fm["L"]) == L:
            for j, p in enumerate(fm["requests"]):
                eb.append(rel_err(r["bary"][p], fm["bary"][j]))
                el.append(rel_err(r["grad_lambda"][p], fm["grad_lambda"][j]))
                hist.append(int(p))
        err = {"bary": max(eb), "grad_lambda": max(el),
               "grad_lambda_per_histogram": [float("%.3g" % e) for e in el],
               "grad_dic": rel_err(gd[0], fx["grad_dic"]), "histograms_checked": hist,
               "against": "reference compute_gradient_both_log_signArray (FP64, L=200, "
                          "tests/golden/large_c3*.npz; dL/datoms of histogram 0)"}
    footprint = wdl.SinkhornSolver.footprint(n, n, K, P, L, wdl.FLAG_LOG_DOMAIN)
    s.close()
    e0.destroy()
    e1.destroy()
    out = {"workload": "log domain, small epsilon: 256x256, 32 curve histograms, K=6, "
                       "gamma=0.5, L=200, KL loss, all gradients",
           "steps": steps, "ms_per_step": round(ms, 3),
           "value": round(P * L / (ms / 1e3), 1), "unit": "Sinkhorn-iters/s",
           "log_range_events": events, "hbm_footprint_gb": round(footprint / 1e9, 2),
           "tuning": tuning,
           "err_vs_fp64": err}
    out.update(stage)
    return out


def host_cpus():
    """The host's CPU as lscpu describes it, and the physical cores this
    process may run on (its affinity mask, one per core however many
    hardware threads it has)."""
    info = {}
    try:
        txt = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in txt.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in ("Model name", "Socket(s)", "Core(s) per socket", "Thread(s) per core",
                             "CPU(s)"):
                info[k.strip()] = v.strip()
    except (OSError, subprocess.SubprocessError):
        pass
    cpus = sorted(os.sched_getaffinity(0))
    cores = set()
    for cpu in cpus:
        try:
            base = "/sys/devices/system/cpu/cpu%d/topology/" % cpu
            with open(base + "physical_package_id") as f1, open(base + "core_id") as f2:
                cores.add((f1.read().strip(), f2.read().strip()))
        except OSError:
            cores.add(("?", str(cpu)))
    info["visible_cpus"] = len(cpus)
    info["visible_physical_cores"] = len(cores)
    return info


def cpu_baseline(c, gpu, budget_s=15.0):
    """Reference CPU path (oracle/_ref, compiled from the reference sources) on
    a bounded sample of the same workload, and the GPU error against it.
    OpenMP runs one thread per visible physical core (at most the process's
    OMP_NUM_THREADS), threads bound close, one place per core."""
    hc = host_cpus()
    nthreads = hc["visible_physical_cores"]
    if os.environ.get("OMP_NUM_THREADS", "").isdigit():
        nthreads = min(nthreads, int(os.environ["OMP_NUM_THREADS"]))
    # read by the OpenMP runtime when oracle/_ref loads it, below
    os.environ["OMP_NUM_THREADS"] = str(nthreads)
    os.environ["OMP_PROC_BIND"] = "close"
    os.environ["OMP_PLACES"] = "cores"
    from oracle import ref
    if not ref.available():
        return None, None, []
    n, K, L, g, loss = c["n"], c["K"], c["L"], c["gamma"], c["loss"]
    out = {"cores": ref.num_threads(), "kind": "reference", "omp": "OMP_PROC_BIND=close "
           "OMP_PLACES=cores OMP_NUM_THREADS=%d" % nthreads, "host": hc}
    errs = {"grad_lambda": [], "grad_dic": [], "bary": []}
    refs = []
    t_sse, done = 0.0, 0
    for p in range(c["P"]):
        t0 = time.perf_counter()
        r = ref.gradient_both(c["atoms"], c["weights"][p], c["obs"][p], n, n, g, L, loss,
                              variant="sse")
        t_sse += time.perf_counter() - t0
        done += 1
        refs.append((p, r))
        for k, gk in (("grad_lambda", gpu["grad_lambda"][p]), ("grad_dic", gpu["grad_dic"][p]),
                      ("bary", gpu["bary"][p])):
            errs[k].append(float(np.linalg.norm(gk.ravel() - r[k].ravel())
                                 / np.linalg.norm(r[k].ravel())))
        if t_sse > budget_s:
            break
    out["value"] = done * L / t_sse
    out["unit"] = "Sinkhorn-iters/s"
    out["sample"] = ("compute_gradient_both with GaussianSSE2DKernel (AVX, OpenMP) on the first %d "
                     "of the %d histograms at full size (128x128, K=5, L=100; histograms until "
                     "%.0f s of CPU time), %.1f s" % (done, c["P"], budget_s, t_sse))
    # the app's default kernel (Gaussian2DKernel, main_dictionary_learning.cpp:176), one histogram
    t0 = time.perf_counter()
    ref.gradient_both(c["atoms"], c["weights"][0], c["obs"][0], n, n, g, L, loss, variant="plain")
    t_plain = time.perf_counter() - t0
    out["plain_kernel_value"] = L / t_plain
    # one OpenMP thread, the SSE/AVX kernel, one histogram
    ref.set_threads(1)
    t0 = time.perf_counter()
    ref.gradient_both(c["atoms"], c["weights"][0], c["obs"][0], n, n, g, L, loss, variant="sse")
    out["one_thread_value"] = L / (time.perf_counter() - t0)
    ref.set_threads(out["cores"])
    out["cpu_model"] = hc.get("Model name") or cpu_model()
    # the log domain (LogSignArrayGaussian2DKernel) on the 256x256 log workload:
    # 2 Sinkhorn iterations of one histogram (its O(n^3) exact passes take ~minutes
    # per histogram at L=200), as iterations/s per histogram
    cl = synthetic.make_problem("log_small_eps", P=1)
    t0 = time.perf_counter()
    ref.gradient_both(cl["atoms"], cl["weights"][0], cl["obs"][0], cl["n"], cl["n"], cl["gamma"], 2,
                      cl["loss"], variant="log")
    out["log_domain"] = {"value": round(2 / (time.perf_counter() - t0), 3),
                         "unit": "Sinkhorn-iters/s",
                         "sample": "compute_gradient_both_log_signArray, 256x256, K=6, gamma=0.5, "
                                   "KL, one curve histogram, 2 iterations (forward + backward)"}
    # the projection workload (128x128, K=8 blob dictionary, gamma=1) and the
    # served workload's 64x64 and 128x128 dictionaries: the reference's
    # log-domain solver, which computes both chains (a projection needs the
    # weight chain only), a few iterations of one request each
    cp = synthetic.make_problem("projection", P=1)
    atoms_p = synthetic.blob_histograms(np.random.default_rng(7), cp["n"], cp["K"], cp["nblobs"],
                                        cp["sigma"])
    t0 = time.perf_counter()
    ref.gradient_both(atoms_p, cp["weights"][0], cp["obs"][0], cp["n"], cp["n"], cp["gamma"], 4,
                      cp["loss"], variant="log")
    out["projection"] = {"value": round(4 / (time.perf_counter() - t0), 3),
                         "unit": "Sinkhorn-iters/s",
                         "sample": "compute_gradient_both_log_signArray (both gradient chains), "
                                   "one projection request, 128x128, K=8, gamma=1, 4 iterations"}
    wl, arrivals, _ = synthetic.served_workload()
    served = {}
    for ns, Ls in ((64, 20), (128, 4)):
        Ks, gs, atoms_s = wl[ns]
        h = next(a[3] for a in arrivals if a[1] == ns)
        t0 = time.perf_counter()
        ref.gradient_both(atoms_s, np.full(Ks, 1.0 / Ks), h, ns, ns, gs, Ls, 2, variant="log")
        served[str(ns)] = {"value": round(Ls / (time.perf_counter() - t0), 3),
                           "unit": "Sinkhorn-iters/s",
                           "sample": "one served request, %dx%d, K=%d, gamma=%g, %d iterations"
                                     % (ns, ns, Ks, gs, Ls)}
    out["served"] = served
    # Gaussian3DKernel::convolve (kernels.h:762-807) on one 64^3 volume
    v = np.random.default_rng(1).uniform(0, 1, LINE_3D[0] ** 3)
    t0 = time.perf_counter()
    ref.convolve3d(v, LINE_3D[0], LINE_3D[0], LINE_3D[0], LINE_3D[1])
    out["conv3d"] = {"value": round(1.0 / (time.perf_counter() - t0), 2), "unit": "volumes/s",
                     "sample": "Gaussian3DKernel::convolve, one 64^3 volume, gamma=4"}
    err = {k: max(v) for k, v in errs.items()}
    err["histograms_checked"] = done
    err["against"] = "reference compute_gradient_both (FP64, CPU)"
    return out, err, refs


LINE_3D = (64, 4.0, 16)     # side, gamma, volumes per call
LINE_1D = (1024, 25.0, 4096)  # bins, gamma, vectors per call


def line_leg(ctx, reps=20):
    """The 1-D and 3-D KernelTypes (Gaussian1DKernel, Gaussian3DKernel,
    kernels.h:501-815), FP32 device forms, inputs resident in HBM.  Algorithmic
    FLOPs per the dense separable count: 2 n len per pass of n values along
    lines of len (2 W^2 per 1-D vector; 2 N (W + H + D) per volume)."""
    out = {}
    rng = np.random.default_rng(2)
    e0, e1 = _hip.Event(), _hip.Event()
    n3, g3, v3 = LINE_3D
    n1, g1, v1 = LINE_1D
    for name, k, n, nv, flops in (
            ("conv3d", wdl.HipGaussian3DKernel(ctx, g3, n3, n3, n3), n3 ** 3, v3,
             2.0 * n3 ** 3 * 3 * n3),
            ("conv1d", wdl.HipGaussian1DKernel(ctx, g1, n1), n1, v1, 2.0 * n1 * n1)):
        din = _hip.DeviceArray((nv * n,))
        dout = _hip.DeviceArray((nv * n,))
        din.upload(rng.uniform(0, 1, nv * n))
        k.convolve_device(din.ptr, dout.ptr, nv)
        e0.record(ctx.stream)
        for _ in range(reps):
            k.convolve_device(din.ptr, dout.ptr, nv)
        e1.record(ctx.stream)
        e1.synchronize()
        ms = e0.elapsed_ms(e1) / reps
        out[name] = {"ms_per_call": round(ms, 4), "vectors_per_call": nv,
                     "value": round(nv / (ms * 1e-3), 1),
                     "unit": "volumes/s" if name == "conv3d" else "vectors/s",
                     "dense_tflops": round(nv * flops / (ms * 1e-3) / 1e12, 2)}
        din.free()
        dout.free()
        k.close()
    out["conv3d"]["shape"] = "%d^3, gamma=%g" % (n3, g3)
    out["conv1d"]["shape"] = "%d bins, gamma=%g" % (n1, g1)
    e0.destroy()
    e1.destroy()
    return out


def evaluate_leg(s, c, solve_ms, reps=5):
    """evaluate_both (inverseWasserstein.h:986-1145) around the batched solve:
    the reparametrisation, loss sum, projections and chain rule on the device.
    Timed with x and g resident in HBM (wdl_evaluate_both_device) and as the
    lbfgs callback with x and g in host memory (PCIe-inclusive)."""
    P, K, N = c["P"], c["K"], c["n"] * c["n"]
    scale = (N / P) / c["scale_factor"]
    # x as regress_both builds it (inverseWasserstein.h:389-401): log weights
    # times the dictionary scale, then log atoms
    x = np.concatenate([np.log(np.asarray(c["weights"], np.float64)).ravel() * scale,
                        np.log(np.asarray(c["atoms"], np.float64)).ravel()])
    s.set_scale(scale, True)
    dx = _hip.DeviceArray.from_host(x, np.float64)
    dg = _hip.DeviceArray(x.shape, np.float64)
    s.evaluate_both_device(dx.ptr, dg.ptr)
    t0 = time.perf_counter()
    for _ in range(reps):
        f = s.evaluate_both_device(dx.ptr, dg.ptr)
    dev_ms = (time.perf_counter() - t0) / reps * 1e3
    t0 = time.perf_counter()
    for _ in range(reps):
        fh, _g = s.evaluate_both(x)
    host_ms = (time.perf_counter() - t0) / reps * 1e3
    dx.free()
    dg.free()
    return {"device_ms": round(dev_ms, 3), "callback_ms": round(host_ms, 3),
            "solve_ms": round(solve_ms, 3), "f": f, "f_callback": fh,
            "pcie_bytes_per_call": int(2 * x.nbytes),
            "what": "one evaluation: reparametrisation + batched solve + loss sum, projections "
                    "and chain rule, all on the device; device_ms with x and g in HBM, "
                    "callback_ms with x and g in host memory"}


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def errors_against(refs, res):
    e = {"grad_lambda": 0.0, "grad_dic": 0.0, "bary": 0.0}
    for p, r in refs:
        for k in e:
            e[k] = max(e[k], float(np.linalg.norm(res[k][p].ravel() - r[k].ravel())
                                   / np.linalg.norm(r[k].ravel())))
    return e


def bf16x3_leg(ctx, c, steps):
    """The same batched workload with the convolution passes in split-BF16x3,
    and the bare convolution in both precisions."""
    n, P, K, L, g, loss = c["n"], c["P"], c["K"], c["L"], c["gamma"], c["loss"]
    s = wdl.SinkhornSolver(ctx, n, n, K, P, L, g, loss, wdl.FLAG_BF16X3)
    s.set_observed(c["obs"])
    s.set_problem(c["atoms"], c["weights"])
    s.autotune(2)
    stream = ctx.stream
    e0, e1 = _hip.Event(), _hip.Event()
    s.run()
    e0.record(stream)
    for _ in range(steps):
        s.run()
    e1.record(stream)
    e1.synchronize()
    ms = e0.elapsed_ms(e1) / steps
    res = s.results()
    res["grad_dic"] = s.grad_dic()
    s.run_stage(wdl.STAGE_BACKWARD, L // 2)
    e0.record(stream)
    for _ in range(20):
        s.run_stage(wdl.STAGE_BACKWARD, L // 2)
    e1.record(stream)
    e1.synchronize()
    bwd_ms = e0.elapsed_ms(e1) / 20
    s.close()
    nimg = 4 * P * K
    dbuf_in = _hip.DeviceArray((nimg, n, n))
    dbuf_out = _hip.DeviceArray((nimg, n, n))
    dbuf_in.upload(np.random.default_rng(0).uniform(0, 1, (nimg, n, n)))
    conv = {}
    for name, f in (("fp32", 0), ("bf16x3", wdl.FLAG_BF16X3)):
        kern = wdl.HipGaussian2DKernel(ctx, g, n, n, f)
        kern.convolve_device(dbuf_in.ptr, dbuf_out.ptr, nimg)
        e0.record(stream)
        for _ in range(20):
            kern.convolve_device(dbuf_in.ptr, dbuf_out.ptr, nimg)
        e1.record(stream)
        e1.synchronize()
        conv[name + "_launch_ms"] = round(e0.elapsed_ms(e1) / 20, 4)
        kern.close()
    dbuf_in.free()
    dbuf_out.free()
    e0.destroy()
    e1.destroy()
    return {"ms_per_step": round(ms, 4), "value": round(P * L / (ms / 1e3), 1),
            "unit": "Sinkhorn-iters/s", "backward_launch_ms": round(bwd_ms, 4),
            "conv_only": conv}, res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain-form timing")
    ap.add_argument("--dense", action="store_true", help="do not skip all-zero factor chunks")
    ap.add_argument("--streams", type=int, default=0,
                    help="histogram groups on separate streams (0: autotuned)")
    ap.add_argument("--no-projection", action="store_true", help="skip the projection leg")
    ap.add_argument("--no-log", action="store_true", help="skip the 256x256 log-domain leg")
    ap.add_argument("--no-served", action="store_true", help="skip the served projection leg")
    ap.add_argument("--no-uneven", action="store_true",
                    help="skip the uneven-grid-sizes (solver group) leg")
    ap.add_argument("--no-bf16x3", action="store_true", help="skip the split-BF16x3 leg")
    ap.add_argument("--no-line", action="store_true", help="skip the 1-D / 3-D kernel leg")
    ap.add_argument("--projection-requests", type=int, default=4096,
                    help="how many of the 4096 projection requests to run")
    ap.add_argument("--no-one-launch", action="store_true",
                    help="skip the projection leg's single checkpointed launch of all requests")
    ap.add_argument("--served-seconds", type=float, default=20.0,
                    help="arrival window of the served leg")
    ap.add_argument("--copy-calib", action="store_true",
                    help="also run one 256 MiB device-to-device copy (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: the library has no multi-GPU path "
              "(the P solves would shard by histogram with one K*N all-reduce per "
              "evaluation). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    c = synthetic.make_problem("batched_std")
    n, P, K, L, g, loss = c["n"], c["P"], c["K"], c["L"], c["gamma"], c["loss"]
    N = n * n
    log("workload batched_std n=%d P=%d K=%d L=%d gamma=%g" % (n, P, K, L, g))
    ctx = wdl.Context(0)
    flags = wdl.FLAG_DENSE if args.dense else 0
    s = wdl.SinkhornSolver(ctx, n, n, K, P, L, g, loss, flags)
    s.set_observed(c["obs"])
    s.set_problem(c["atoms"], c["weights"])
    if args.streams:
        s.set_streams(args.streams)
        nstreams = args.streams
    else:
        nstreams = s.autotune(3)
        log("autotune: %s" % s.tuning())
    stream = ctx.stream

    for _ in range(args.warmup):
        s.run(wdl.RUN_BATCHED)
    ctx.synchronize()
    log("warmup done")

    ev = [_hip.Event() for _ in range(args.steps + 1)]
    _hip.synchronize()
    ev[0].record(stream)
    for k in range(args.steps):
        s.run(wdl.RUN_BATCHED)
        ev[k + 1].record(stream)
    ev[-1].synchronize()
    ctx.synchronize()
    total_ms = ev[0].elapsed_ms(ev[-1])
    step_ms = np.array([ev[k].elapsed_ms(ev[k + 1]) for k in range(args.steps)])
    value = P * L * args.steps / (total_ms / 1e3)
    log("batched: %.3f ms/step, %.0f iters/s" % (total_ms / args.steps, value))

    res = s.results()
    res["grad_dic"] = s.grad_dic()

    plain = None
    if not args.no_plain:
        e0, e1 = _hip.Event(), _hip.Event()
        e0.record(stream)
        s.run(wdl.RUN_PLAIN)
        e1.record(stream)
        e1.synchronize()
        pms = e0.elapsed_ms(e1)
        plain = {"value": P * L / (pms / 1e3), "ms_per_step": pms,
                 "form": "same kernels, one solve per launch sequence, P solves in turn, no graph"}
        log("plain: %.1f ms/step" % pms)

    # Dominant kernel (backward stage: both unrolled chains, 4 image-convolutions
    # per (histogram, atom)) timed alone on the same stream.
    reps = 20
    e0, e1 = _hip.Event(), _hip.Event()
    step = L // 2
    s.run_stage(wdl.STAGE_BACKWARD, step)
    e0.record(stream)
    for _ in range(reps):
        s.run_stage(wdl.STAGE_BACKWARD, step)
    e1.record(stream)
    e1.synchronize()
    bwd_ms = e0.elapsed_ms(e1) / reps
    e0.record(stream)
    for _ in range(reps):
        s.run_stage(wdl.STAGE_FORWARD, step)
    e1.record(stream)
    e1.synchronize()
    fwd_ms = e0.elapsed_ms(e1) / reps
    kern = wdl.HipGaussian2DKernel(ctx, g, n, n, wdl.FLAG_DENSE if args.dense else 0)
    D = 1 << 20 if args.dense else band_halfwidth(kern.kernel1d())
    f_exec, f_dense = conv_flops(n, D)
    # The bare convolution kernel (no fused elementwise work) on one backward
    # stage's worth of images, for the MFMA efficiency of the conv alone.
    nimg = 4 * P * K
    dbuf_in = _hip.DeviceArray((nimg, n, n))
    dbuf_out = _hip.DeviceArray((nimg, n, n))
    dbuf_in.upload(np.random.default_rng(0).uniform(0, 1, (nimg, n, n)))
    kern.convolve_device(dbuf_in.ptr, dbuf_out.ptr, nimg)
    e0.record(stream)
    for _ in range(reps):
        kern.convolve_device(dbuf_in.ptr, dbuf_out.ptr, nimg)
    e1.record(stream)
    e1.synchronize()
    conv_ms = e0.elapsed_ms(e1) / reps
    dbuf_in.free()
    dbuf_out.free()
    convs_bwd = 4 * P * K
    # algorithmic FLOPs: 4n^3 per image-convolution (kernels.h:680-692), the
    # dense count the roofline is declared on; the banded count the MFMAs
    # execute (zero-factor chunks skipped) is reported beside it
    achieved = convs_bwd * f_dense / (bwd_ms * 1e-3) / 1e12
    executed = convs_bwd * f_exec / (bwd_ms * 1e-3) / 1e12
    traffic, traffic_src = measured_traffic()
    # algorithmic HBM bytes of one backward launch (FP32), each distinct array
    # once: per (p,i) reads phi^s, c^{s-1}, phi^{s-1}, v_i, r_i, resultDic,
    # plus q^s and q^{s-1} shared by the K atoms, and writes resultDic, v, r
    # (the K-fold v / r joins and the atoms are L2 re-reads).
    bytes_bwd = int(P * N * 4 * (9 * K + 2))
    roofline = {
        "bound": "mfma", "achieved": round(achieved, 2), "peak": PEAK_FP32_MFMA,
        "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP32_MFMA, 4), "traffic": traffic,
        "traffic_source": traffic_src,
        "kernel": "k_backward (both unrolled chains, one Sinkhorn step, all 64 histograms)",
        "launch_ms": round(bwd_ms, 4), "forward_launch_ms": round(fwd_ms, 4),
        "flops_per_launch": convs_bwd * f_dense,
        "flop_count": "algorithmic: 4n^3 per image-convolution, 4 per (histogram, atom)",
        "executed_mfma_flops_per_launch": convs_bwd * f_exec,
        "executed_mfma_tflops": round(executed, 2),
        "executed_frac": round(executed / PEAK_FP32_MFMA, 4),
        "hbm_frac": round((traffic if traffic else bytes_bwd) / (bwd_ms * 1e-3) / 1e9 / PEAK_HBM, 4),
        "hbm_frac_basis": "measured traffic" if traffic else "algorithmic bytes",
        "hbm_algorithmic_bytes_per_launch": bytes_bwd,
        "hbm_achieved_gbs": round(bytes_bwd / (bwd_ms * 1e-3) / 1e9, 1),
        "band_halfwidth": D,
        "conv_only": {"images": nimg, "launch_ms": round(conv_ms, 4),
                      "executed_tflops": round(nimg * f_exec / (conv_ms * 1e-3) / 1e12, 2)},
    }
    log("backward launch %.4f ms -> %.1f TFLOP/s algorithmic, %.1f executed"
        % (bwd_ms, achieved, executed))
    whole_flops = P * (6 * L - 3) * K * f_dense
    line = {
        "metric": METRIC, "value": round(value, 1), "unit": "Sinkhorn-iters/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(total_ms / args.steps, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic: blob histograms (1-4 blobs, sigma 4-16 px, floor 1e-6), "
                "uniform random atoms and weights, one seed",
        "config": {"workload": "batched standard domain: compute_gradient_both for all P "
                               "histograms per step", "n": n, "P": P, "K": K, "L": L,
                   "gamma": g, "loss": "quadratic", "band_skip": not args.dense,
                   "streams": nstreams, "tuning": s.tuning(),
                   "tuned_by": "wdl_solver_autotune" if not args.streams else "--streams"},
        "latency_ms": {"p50": round(float(np.percentile(step_ms, 50)), 4),
                       "p99": round(float(np.percentile(step_ms, 99)), 4),
                       "what": "per-step (whole batch) latency"},
        "dense_4n3_tflops_whole_step": round(whole_flops / (total_ms / args.steps * 1e-3) / 1e12, 2),
        "plain": plain,
        "batched_over_plain": (round(value / plain["value"], 2) if plain else None),
        "roofline": roofline,
    }
    line["evaluate_both"] = evaluate_leg(s, c, total_ms / args.steps)
    log("evaluate_both: %s" % json.dumps(line["evaluate_both"]))
    if not args.no_line:
        line["line_kernels"] = line_leg(ctx)
        log("1-D / 3-D kernels: %s" % json.dumps(line["line_kernels"]))
    if not args.no_projection and args.projection_requests > 0:
        line["projection"] = projection_leg(ctx, args.projection_requests,
                                            one_launch=not args.no_one_launch)
        log("projection: %s iters/s" % line["projection"]["value"])
    if not args.no_log:
        line["log_small_eps"] = log_leg(ctx)
        log("log domain 256x256: %s iters/s" % line["log_small_eps"]["value"])
    if not args.no_served:
        line["served"] = served_leg(args.served_seconds)
        log("served: %s" % json.dumps(line["served"]["latency_ms"]))
        line["served_mixed"] = served_leg(args.served_seconds, mixed=True)
        log("served, mixed batching: %s" % json.dumps(line["served_mixed"]["latency_ms"]))
    if not args.no_uneven:
        line["uneven_sizes"] = uneven_leg()
        log("uneven sizes: group %.1f ms, streams %.1f ms per round" %
            (line["uneven_sizes"]["group"]["ms_per_round"],
             line["uneven_sizes"]["streams"]["ms_per_round"]))
    bres = None
    if not args.no_bf16x3:
        line["bf16x3"], bres = bf16x3_leg(ctx, c, args.steps)
        log("bf16x3: %s iters/s" % line["bf16x3"]["value"])
    if not args.no_cpu:
        cb, err, refs = cpu_baseline(c, res)
        line["cpu_baseline"] = cb
        line["err_vs_fp64"] = err
        if bres is not None and refs:
            line["bf16x3"]["err_vs_fp64"] = errors_against(refs, bres)
    if args.copy_calib:
        # A plain copy of known size: FETCH_SIZE / WRITE_SIZE of this copy kernel
        # scale the counters read for the solver kernels to bytes.
        nb = 256 << 20
        src = _hip.DeviceArray((nb // 4,))
        dst = _hip.DeviceArray((nb // 4,))
        _hip.rt().hipMemcpy(dst.ptr, src.ptr, nb, _hip.D2D)
        _hip.synchronize()
        line["copy_calibration_bytes"] = nb
        src.free()
        dst.free()
    for e in ev + [e0, e1]:
        e.destroy()
    ctx.close()  # closes the solver and the kernel object first
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
